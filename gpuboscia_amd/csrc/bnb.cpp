// Host branch-and-bound driver behind gb_solve.
//
// Follows gpuboscia::solve / DeviceResidentHybridSolver
// (gpuboscia.cuh:56-170, device_resident_solver.cuh:1184-2085) with one
// deliberate change of schedule: every node of a completed batch is
// post-processed in batch order, on one host thread.  The plain form keeps one
// batch in flight, so the GPU run and the CPU oracle expand the same tree; the
// batched forms keep exactly two (the next batch is popped and launched before
// the current one is post-processed).  The reference overlaps up to
// num_streams batches on host threads, which makes node counts and
// tie-breaking schedule dependent; a fixed alternation keeps the search
// deterministic.  The frontier is a best-first heap ordered by (dual_bound, id).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <stdexcept>
#include <string>
#include <vector>

#include "gb_fast.h"
#include "gb_host.h"
#include "gb_mlp.h"
#include "gb_qap32.h"

// Children a BranchStrategy callback returns (gb_children_push).
struct gb_children {
    struct Child {
        int id, parent, depth;
        double bound;
        std::vector<double> lb, ub;
    };
    int n = 0;
    std::vector<Child> kids;
};

namespace gb {
namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

// Node bounds and warm starts live in a chunked slab (3 n doubles per slot:
// lb, ub, relaxation_x); heap entries are small records.  This keeps host
// post-processing at a few hundred ns per node (no per-child allocation).
// A slot is reference counted: children that warm-start from their parent's
// iterate hold the parent's slot (whose relaxation_x is that iterate) until
// they are staged.
// Slots of node data.  Each holds the node's bounds and its relaxation x
// (n doubles).  The bounds are 2n doubles, or, in compact mode (a problem
// whose bounds are all 0 or 1 and whose branching rounds fractional values
// only, so every node bound stays 0 or 1), 2n bytes: child creation and
// staging then move a byte per bound instead of eight.
struct NodePool {
    int n = 0;
    bool compact = false;
    static constexpr int kChunk = 4096;
    std::vector<std::unique_ptr<double[]>> chunks;
    std::vector<std::unique_ptr<uint8_t[]>> bchunks;  // compact: lb, ub bytes
    std::vector<int> free_slots;
    std::vector<int> refs;
    int next = 0;
    // device iterates: relaxation_x of slot s lives in the device store
    // (dchunks) instead of a host chunk
    const std::vector<double*>* dchunks = nullptr;
    int alloc() {
        int s;
        if (!free_slots.empty()) {
            s = free_slots.back();
            free_slots.pop_back();
        } else {
            if (next % kChunk == 0) {
                if (!(compact && dchunks)) chunks.emplace_back(new double[(size_t)kChunk * (compact ? 1 : 3) * n]);
                if (compact) bchunks.emplace_back(new uint8_t[(size_t)kChunk * 2 * n]);
            }
            s = next++;
            refs.push_back(0);
        }
        refs[s] = 1;
        return s;
    }
    void retain(int slot) { ++refs[slot]; }
    void release(int slot) {
        if (--refs[slot] == 0) free_slots.push_back(slot);
    }
    // double bounds (not compact)
    double* lb(int slot) { return chunks[slot / kChunk].get() + (size_t)(slot % kChunk) * 3 * n; }
    double* ub(int slot) { return lb(slot) + n; }
    double* rx(int slot) {
        return compact ? chunks[slot / kChunk].get() + (size_t)(slot % kChunk) * n : lb(slot) + 2 * n;
    }
    // device address of slot's iterate (device iterates only)
    unsigned long long drx(int slot) const {
        return reinterpret_cast<unsigned long long>((*dchunks)[slot / kChunk] + (size_t)(slot % kChunk) * n);
    }
    // byte bounds (compact): lb8, then ub8
    uint8_t* lb8(int slot) { return bchunks[slot / kChunk].get() + (size_t)(slot % kChunk) * 2 * n; }
    uint8_t* ub8(int slot) { return lb8(slot) + n; }
};

struct HostNode {
    int id = 0, parent = -1, depth = 0;
    int slot = -1;                   // NodePool slot holding lb, ub, relaxation_x
    bool has_rx = false;             // warm start present (relaxation_x non-empty)
    int xslot = -1;                  // or: clip the relaxation_x of this slot (the parent's) to the box
    int riters = 0;                  // relaxation_iterations
    double bound = -1e300;           // dual_bound
    // single-variable branching record for pseudocosts
    int bvar = -1;                   // branched variable (-1: none / multi-variable)
    int bdir = 0;                    // 0 = down (ub = floor), 1 = up (lb = ceil)
    double bfrac = 0.0;              // distance moved by the branching bound
    double pbound = 0.0;             // parent's certified bound
};

// A node pushed on a solver object (gb_solver_push), with its own bounds and
// iterate.
struct InitNode {
    int id = 0, parent = -1, depth = 0;
    std::vector<double> lb, ub, rx;
    int riters = 0;
    double bound = -1e300;
};

// What a solver object adds to a solve: its pushed nodes, its scheduling
// observer, and the accessors it reports (device_resident_solver.cuh:1280-1298).
struct SolveHooks {
    const std::vector<InitNode>* nodes = nullptr;
    gb_scheduling_fn sched = nullptr;
    void* sched_user = nullptr;
    int last_iters = 0;
    double last_gap = std::numeric_limits<double>::infinity();
    bool last_conv = false;
};

// Pseudocosts: mean certified-bound gain per unit of fractionality removed,
// per variable and direction, from the children relaxed so far.
struct Pseudocosts {
    std::vector<double> up_sum, dn_sum;
    std::vector<int> up_cnt, dn_cnt;
    double up_all = 0.0, dn_all = 0.0;
    int up_n = 0, dn_n = 0;
    void init(int n) {
        up_sum.assign(n, 0.0); dn_sum.assign(n, 0.0);
        up_cnt.assign(n, 0); dn_cnt.assign(n, 0);
    }
    void record(int var, int dir, double gain_per_unit) {
        if (dir) { up_sum[var] += gain_per_unit; ++up_cnt[var]; up_all += gain_per_unit; ++up_n; }
        else { dn_sum[var] += gain_per_unit; ++dn_cnt[var]; dn_all += gain_per_unit; ++dn_n; }
    }
    // relative pseudocost: variable mean / mean over all records (1 without history)
    double up(int v) const {
        if (up_n == 0) return 1.0;
        const double mean = up_all / up_n;
        const double pv = up_cnt[v] ? up_sum[v] / up_cnt[v] : mean;
        return pv / (mean + 1e-12);
    }
    double dn(int v) const {
        if (dn_n == 0) return 1.0;
        const double mean = dn_all / dn_n;
        const double pv = dn_cnt[v] ? dn_sum[v] / dn_cnt[v] : mean;
        return pv / (mean + 1e-12);
    }
};

// The 8 scorer features of candidate variables (order documented in
// gpuboscia_amd.h next to GB_MLP_FEATURES).  Features 3 and 4 are the
// estimated bound gains of the up and down branches: pseudocost (mean gain
// per unit from relaxed children) times the distance moved, or, without
// history for the variable, the quadratic's own estimate 1/2 s_g Q_ii d^2;
// each is divided by its mean over the node's candidates.
void candidate_features(const double* x, const double* g, const double* lb, const double* ub, int n,
                        int depth, double gap, double fval, const Pseudocosts& pc,
                        const std::vector<double>& qdiag, double gs, const std::vector<int>& cands,
                        float* out) {
    double gmax = 0.0;
    for (int j = 0; j < n; ++j) gmax = std::max(gmax, std::fabs(g[j]));
    const double rgap = gap / (1.0 + std::fabs(fval));
    std::vector<double> gu(cands.size()), gd(cands.size());
    double su = 0.0, sd = 0.0;
    for (size_t c = 0; c < cands.size(); ++c) {
        const int v = cands[c];
        const double du = std::ceil(x[v]) - x[v], dd = x[v] - std::floor(x[v]);
        const double curv = 0.5 * gs * (qdiag.empty() ? 0.0 : qdiag[v]);
        gu[c] = (pc.up_cnt[v] ? pc.up_sum[v] / pc.up_cnt[v] : curv * du) * du;
        gd[c] = (pc.dn_cnt[v] ? pc.dn_sum[v] / pc.dn_cnt[v] : curv * dd) * dd;
        su += gu[c];
        sd += gd[c];
    }
    const double mu = su / (cands.empty() ? 1.0 : (double)cands.size());
    const double md = sd / (cands.empty() ? 1.0 : (double)cands.size());
    for (size_t c = 0; c < cands.size(); ++c) {
        const int v = cands[c];
        const double frac = x[v] - std::floor(x[v]);
        float* f = out + c * kMlpFeatures;
        f[0] = (float)x[v];
        f[1] = (float)std::min(frac, 1.0 - frac);
        f[2] = (float)(g[v] / (gmax + 1e-12));
        f[3] = (float)(gu[c] / (mu + 1e-12));
        f[4] = (float)(gd[c] / (md + 1e-12));
        f[5] = (float)depth / (float)n;
        f[6] = (float)rgap;
        f[7] = (float)(ub[v] - lb[v]);
    }
}

std::vector<int> fractional_candidates(const double* x, const std::vector<int>& iv, double tol) {
    std::vector<int> c;
    for (int v : iv) {
        const double frac = x[v] - std::floor(x[v]);
        if (std::min(frac, 1.0 - frac) > tol) c.push_back(v);
    }
    return c;
}

// Strong-branching label collection (gb_branching_dataset).
struct Labeler {
    int max_nodes = 0;
    float* features = nullptr;
    float* labels = nullptr;
    int capacity = 0;
    int count = 0;
    int labeled = 0;
};

// Best-first frontier ordered by (dual_bound, id).  (bound, id) is a strict
// order (ids are unique), so nodes leave in exactly that order whatever the
// structure.  Entries are families: nodes with one bound and consecutive ids
// (the children of one branching), which leave one by one; no other node's
// key can fall between two members of a family, so a family stays the
// smallest entry until its last member has left, and costs one heap push and
// one pop.  A 4-ary heap of 16-byte keys holds every key below a threshold
// (at most a few ten thousand, cache resident), an unsorted vector the rest;
// when the heap runs empty the smallest keys of the vector are selected
// (nth_element) into it, and when it grows too large its largest keys move to
// the vector.  Family records live in recycled blocks.
class Frontier {
  public:
    bool empty() const { return hot_.empty(); }
    size_t size() const { return nodes_; }
    const HostNode& top() const {
        const Family& f = fams_[hot_.front().fam];
        return recs_[f.start + f.pos];
    }
    double top_bound() const { return hot_.front().bound; }  // without touching the record
    void push(const HostNode& nd) { push_family(&nd, 1); }
    // nds[0 .. count): consecutive runs with one bound and ids nds[0].id + i
    // become one entry each.
    void push_family(const HostNode* nds, int count) {
        while (count > 0) {
            int run = 1;
            while (run < count && nds[run].bound == nds[0].bound && nds[run].id == nds[0].id + run) ++run;
            push_run(nds, run);
            nds += run;
            count -= run;
        }
    }
    void pop() {
        Key& k = hot_.front();
        Family& f = fams_[k.fam];
        --nodes_;
        if (++f.pos < f.count) {
            k.id = recs_[f.start + f.pos].id;  // the next member: still the smallest key
            return;
        }
        free_blocks_[f.count].push_back(f.start);
        free_fams_.push_back(k.fam);
        heap_pop();
        if (hot_.empty() && !cold_.empty()) refill();
    }

  private:
    static constexpr size_t kHot = 1 << 15;
    struct Key {
        double bound;
        int id, fam;
    };
    struct Family {
        int start = 0, count = 0, pos = 0;
    };
    static bool before(const Key& a, const Key& b) {
        return a.bound < b.bound || (a.bound == b.bound && a.id < b.id);
    }
    void push_run(const HostNode* nds, int count) {
        int f;
        if (!free_fams_.empty()) {
            f = free_fams_.back();
            free_fams_.pop_back();
        } else {
            f = (int)fams_.size();
            fams_.emplace_back();
        }
        if ((size_t)count >= free_blocks_.size()) free_blocks_.resize(count + 1);
        int at;
        if (!free_blocks_[count].empty()) {
            at = free_blocks_[count].back();
            free_blocks_[count].pop_back();
        } else {
            at = (int)recs_.size();
            recs_.resize(recs_.size() + count);
        }
        std::copy(nds, nds + count, recs_.begin() + at);
        fams_[f] = Family{at, count, 0};
        nodes_ += count;
        const Key key{nds[0].bound, nds[0].id, f};
        if (!cold_.empty() && !before(key, lim_)) {
            cold_.push_back(key);
            return;
        }
        heap_push(key);
        if (hot_.size() > 2 * kHot) spill();
    }
    void heap_push(const Key& key) {
        size_t i = hot_.size();
        hot_.push_back(key);
        while (i > 0) {
            const size_t parent = (i - 1) / 4;
            if (!before(key, hot_[parent])) break;
            hot_[i] = hot_[parent];
            i = parent;
        }
        hot_[i] = key;
    }
    void heap_pop() {
        const Key last = hot_.back();
        hot_.pop_back();
        const size_t m = hot_.size();
        if (m == 0) return;
        size_t i = 0;
        for (;;) {
            const size_t c0 = 4 * i + 1;
            if (c0 >= m) break;
            size_t best = c0;
            const size_t ce = std::min(c0 + 4, m);
            for (size_t c = c0 + 1; c < ce; ++c)
                if (before(hot_[c], hot_[best])) best = c;
            if (!before(hot_[best], last)) break;
            hot_[i] = hot_[best];
            i = best;
        }
        hot_[i] = last;
    }
    void sift_down(size_t i) {
        const size_t m = hot_.size();
        const Key key = hot_[i];
        for (;;) {
            const size_t c0 = 4 * i + 1;
            if (c0 >= m) break;
            size_t best = c0;
            const size_t ce = std::min(c0 + 4, m);
            for (size_t c = c0 + 1; c < ce; ++c)
                if (before(hot_[c], hot_[best])) best = c;
            if (!before(hot_[best], key)) break;
            hot_[i] = hot_[best];
            i = best;
        }
        hot_[i] = key;
    }
    void heapify() {  // bottom-up, O(size)
        if (hot_.size() < 2) return;
        for (size_t i = (hot_.size() - 2) / 4 + 1; i-- > 0;) sift_down(i);
    }
    // Keep the kHot smallest keys in the heap, the rest in the vector.
    void spill() {
        std::nth_element(hot_.begin(), hot_.begin() + kHot, hot_.end(), before);
        lim_ = *std::min_element(hot_.begin() + kHot, hot_.end(), before);
        cold_.insert(cold_.end(), hot_.begin() + kHot, hot_.end());
        hot_.resize(kHot);
        heapify();
    }
    // Move the kHot smallest keys of the vector into the (empty) heap.
    void refill() {
        if (cold_.size() > kHot) {
            std::nth_element(cold_.begin(), cold_.begin() + kHot, cold_.end(), before);
            hot_.assign(cold_.begin(), cold_.begin() + kHot);
            cold_.erase(cold_.begin(), cold_.begin() + kHot);
            lim_ = *std::min_element(cold_.begin(), cold_.end(), before);
        } else {
            hot_.swap(cold_);
            cold_.clear();
        }
        heapify();
    }
    std::vector<Key> hot_, cold_;
    Key lim_{0.0, 0, 0};  // every heap key is before lim_, no vector key is (when the vector is non-empty)
    std::vector<HostNode> recs_;
    std::vector<Family> fams_;
    std::vector<int> free_fams_;
    std::vector<std::vector<int>> free_blocks_;  // record blocks by size
    size_t nodes_ = 0;
};

// A fixed set of host worker threads for the per-node loops whose iterations
// are independent (staging copies, branching scans, child bounds): each call
// splits [0, count) into contiguous parts, the caller runs part 0 and waits
// for the rest.  Results do not depend on the thread count.  One call at a
// time: a concurrent caller runs its loop inline.  Created on first use and
// kept for the life of the process.
class HostThreads {
  public:
    static HostThreads& get() {
        static HostThreads* p = new HostThreads();
        return *p;
    }
    template <typename F>
    void parallel_for(int count, F&& f) {
        const int parts = std::min((int)workers_.size() + 1, count / kMinPerPart);
        std::unique_lock<std::mutex> call(call_mu_, std::try_to_lock);
        if (parts <= 1 || !call.owns_lock()) {
            for (int i = 0; i < count; ++i) f(i);
            return;
        }
        std::function<void(int)> body = [&](int p) {
            const int lo = (int)((long long)count * p / parts);
            const int hi = (int)((long long)count * (p + 1) / parts);
            for (int i = lo; i < hi; ++i) f(i);
        };
        {
            std::lock_guard<std::mutex> lk(mu_);
            body_ = &body;
            parts_ = parts;
            next_ = 1;
            left_ = parts - 1;
        }
        cv_.notify_all();
        body(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return left_ == 0; });
        body_ = nullptr;
    }

  private:
    static constexpr int kMinPerPart = 64;
    HostThreads() {
        const unsigned hw = std::thread::hardware_concurrency();
        const int nw = hw > 1 ? (int)std::min(15u, hw - 1) : 0;
        for (int t = 0; t < nw; ++t) std::thread([this] { loop(); }).detach();
        workers_.resize(nw);
    }
    void loop() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return body_ != nullptr && next_ < parts_; });
            const int p = next_++;
            std::function<void(int)>* b = body_;
            lk.unlock();
            (*b)(p);
            lk.lock();
            if (--left_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<char> workers_;  // one entry per worker thread
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    std::function<void(int)>* body_ = nullptr;
    int parts_ = 0, next_ = 0, left_ = 0;
};

// gap_utils.h:14-62
double certified_fw_lower_bound(double inherited, double primal, double gap) {
    if (!std::isfinite(primal) || !std::isfinite(gap) || gap < 0.0) return inherited;
    return std::max(inherited, primal - gap);
}
double gap_absolute(double bound, double inc) {
    if (inc >= kInf || bound <= -kInf) return kInf;
    return std::fabs(inc - bound);
}
double gap_relative(double bound, double inc) {
    if (!std::isfinite(bound) || !std::isfinite(inc)) return kInf;
    const double scale = std::min(std::fabs(bound), std::fabs(inc));
    if (scale == 0.0) return kInf;
    return (inc - bound) / scale;
}
bool gap_within_tolerance(double bound, double inc, double at, double rt) {
    if (!std::isfinite(bound) || !std::isfinite(inc)) return false;
    return gap_absolute(bound, inc) <= at || gap_relative(bound, inc) <= rt;
}
// top_k_infeasible (branching.h:43-65).  Same candidate vector and
// comparator as the reference so std::sort orders ties identically.
std::vector<int> top_k_infeasible(const double* x, const std::vector<int>& iv, int k, double tol) {
    struct Cand { int idx; double dist; };
    std::vector<Cand> c;
    c.reserve(iv.size());
    for (int i : iv) {
        const double frac = x[i] - std::floor(x[i]);
        const double dist = std::min(frac, 1.0 - frac);
        if (dist > tol) c.push_back({i, dist});
    }
    std::sort(c.begin(), c.end(), [](const Cand& a, const Cand& b) { return a.dist > b.dist; });
    const int m = std::min(k, (int)c.size());
    std::vector<int> r(m);
    for (int i = 0; i < m; ++i) r[i] = c[i].idx;
    return r;
}

// top_k_infeasible with the same result, without the allocation and the
// full sort when the k largest distances are distinct and none ties the k-th
// from below (then the sorted prefix is unique); otherwise the full sort of
// top_k_infeasible decides the order of equal distances.  `cand` is scratch.
int top_k_infeasible_into(const double* x, const std::vector<int>& iv, int k, double tol, int* out,
                          std::vector<std::pair<int, double>>& cand) {
    cand.clear();
    int best_i[16];
    double best_d[16];
    int m = 0;
    for (int i : iv) {
        const double frac = x[i] - std::floor(x[i]);
        const double dist = std::min(frac, 1.0 - frac);
        if (!(dist > tol)) continue;
        cand.push_back({i, dist});
        if (m < k || dist > best_d[m - 1]) {
            int p = m < k ? m++ : m - 1;
            while (p > 0 && best_d[p - 1] < dist) {
                best_d[p] = best_d[p - 1];
                best_i[p] = best_i[p - 1];
                --p;
            }
            best_d[p] = dist;
            best_i[p] = i;
        }
    }
    bool tie = false;
    for (int c = 1; c < m && !tie; ++c) tie = best_d[c] == best_d[c - 1];
    if (!tie && m == k)
        for (const auto& cd : cand)
            if (cd.second == best_d[m - 1] && cd.first != best_i[m - 1]) {
                tie = true;
                break;
            }
    if (tie) {
        const std::vector<int> r = top_k_infeasible(x, iv, k, tol);
        for (size_t c = 0; c < r.size(); ++c) out[c] = r[c];
        return (int)r.size();
    }
    for (int c = 0; c < m; ++c) out[c] = best_i[c];
    return m;
}

// Single most fractional variable (branching.h:24-36), ignoring variables
// within the integer tolerance.
int most_fractional(const double* x, const std::vector<int>& iv, double tol) {
    int best = -1;
    double bd = -1.0;
    for (int i : iv) {
        const double frac = x[i] - std::floor(x[i]);
        const double dist = std::min(frac, 1.0 - frac);
        if (dist > tol && dist > bd) { bd = dist; best = i; }
    }
    return best;
}

// Node batch buffers.  The per-node arrays are views into one device block
// and one pinned host block with the same layout, recomputed for each batch
// size N by layout(): inputs first, then the arrays that are both input and
// result, then the results, so a batch moves in one host-to-device copy
// (bounds .. bvv) and one device-to-host copy (x .. topk, and the best
// vertices for small batches; in large ones they stay on the device and the
// host fetches one when it improves the incumbent).  In compact mode (0/1
// bounds, large batches) the bounds travel as bytes and are widened into
// device-only double arrays by a kernel.  With device iterates (xdev) x
// crosses the bus in neither direction: the start points are gathered from
// the device iterate store and the results scattered back into it, and only
// the store addresses travel.
struct DeviceBatch {
    int cap = 0, n = 0;
    bool compact = false;
    bool bv_out = true;  // best vertices in the result copy (else fetched when they improve the incumbent)
    bool xdev = false;   // iterates stay on the device (gathered from and scattered to the iterate store)
    bool learned_io = false;  // depth, pseudocost means and chosen candidates travel with the batch
    int topk_k = 0;           // candidates returned per node
    double *x = nullptr, *lb = nullptr, *ub = nullptr, *bv = nullptr;
    double *eps = nullptr, *gap = nullptr, *val = nullptr, *bvv = nullptr, *cut = nullptr;
    int* it = nullptr;
    uint8_t *st = nullptr, *flags = nullptr;
    double *hx = nullptr, *hlb = nullptr, *hub = nullptr, *hbv = nullptr;
    double *heps = nullptr, *hgap = nullptr, *hval = nullptr, *hbvv = nullptr, *hcut = nullptr;
    int* hit = nullptr;
    uint8_t* hst = nullptr;
    hipStream_t stream = nullptr;  // the stream the batch runs on: its own, or another set's
    hipStream_t own = nullptr;     // the stream this set created (destroyed with it)
    // learned branching / labelling: gradients, feature rows, scores, node
    // depths, per-variable pseudocost means (up, down; < 0 = no history), Q
    // diagonal
    double *g = nullptr, *hg = nullptr;
    float *feat = nullptr, *score = nullptr;
    int *topk = nullptr, *htopk = nullptr;
    int *depth = nullptr, *hdepth = nullptr;
    double *pc = nullptr, *hpc = nullptr;
    double *tkx = nullptr, *htkx = nullptr;   // x at the chosen candidates, [N][k]
    uint8_t *lb8 = nullptr, *hlb8 = nullptr;  // compact: lb bytes then ub bytes, [2][N][n]
    // xdev: per node the store address of its start point and of its iterate
    unsigned long long *xsrc = nullptr, *xdst = nullptr, *hxsrc = nullptr, *hxdst = nullptr;
    unsigned char *dpack = nullptr, *hpack = nullptr;
    size_t in_bytes = 0, out_off = 0, out_bytes = 0;

    struct Offsets {
        size_t lb8, lb, ub, eps, cut, depth, pc, xsrc, xdst, x, it, st, bvv, gap, val, topk, tkx, out_end, bv, dlb,
            dub, end;
    };
    Offsets offsets(int N) const {
        Offsets f{};
        size_t o = 0;
        auto take = [&](size_t bytes) {
            const size_t at = o;
            o = (o + bytes + 255) & ~static_cast<size_t>(255);
            return at;
        };
        const size_t xb = (size_t)N * n * 8;
        f.lb8 = take(compact ? (size_t)2 * N * n : 0);
        f.lb = take(compact ? 0 : xb);
        f.ub = take(compact ? 0 : xb);
        f.eps = take((size_t)N * 8);
        f.cut = take((size_t)N * 8);
        f.depth = take(learned_io ? (size_t)N * 4 : 0);
        f.pc = take(learned_io ? (size_t)n * 16 : 0);
        f.xsrc = take(xdev ? (size_t)N * 8 : 0);
        f.xdst = take(xdev ? (size_t)N * 8 : 0);
        f.x = xdev ? 0 : take(xb);
        f.it = take((size_t)N * 4);
        f.st = take((size_t)N);
        f.bvv = take((size_t)N * 8);
        f.gap = take((size_t)N * 8);
        f.val = take((size_t)N * 8);
        f.topk = take(learned_io ? (size_t)N * topk_k * 4 : 0);
        f.tkx = take(learned_io ? (size_t)N * topk_k * 8 : 0);
        if (bv_out) f.bv = take(xb);
        f.out_end = o;
        // device only
        if (xdev) f.x = take(xb);
        if (!bv_out) f.bv = take(xb);
        f.dlb = take(compact ? xb : 0);
        f.dub = take(compact ? xb : 0);
        f.end = o;
        return f;
    }

    // Point every per-node array at the layout of an N-node batch.
    void layout(int N) {
        const Offsets f = offsets(N);
        auto d = [&](size_t off) { return dpack + off; };
        auto h = [&](size_t off) { return hpack + off; };
        lb = reinterpret_cast<double*>(d(compact ? f.dlb : f.lb));
        ub = reinterpret_cast<double*>(d(compact ? f.dub : f.ub));
        lb8 = compact ? d(f.lb8) : nullptr;
        hlb8 = compact ? h(f.lb8) : nullptr;
        eps = reinterpret_cast<double*>(d(f.eps));
        cut = reinterpret_cast<double*>(d(f.cut));
        x = reinterpret_cast<double*>(d(f.x));
        it = reinterpret_cast<int*>(d(f.it));
        st = d(f.st);
        bvv = reinterpret_cast<double*>(d(f.bvv));
        bv = reinterpret_cast<double*>(d(f.bv));
        gap = reinterpret_cast<double*>(d(f.gap));
        val = reinterpret_cast<double*>(d(f.val));
        hlb = compact ? nullptr : reinterpret_cast<double*>(h(f.lb));
        hub = compact ? nullptr : reinterpret_cast<double*>(h(f.ub));
        heps = reinterpret_cast<double*>(h(f.eps));
        hcut = reinterpret_cast<double*>(h(f.cut));
        hx = xdev ? nullptr : reinterpret_cast<double*>(h(f.x));
        xsrc = xdev ? reinterpret_cast<unsigned long long*>(d(f.xsrc)) : nullptr;
        xdst = xdev ? reinterpret_cast<unsigned long long*>(d(f.xdst)) : nullptr;
        hxsrc = xdev ? reinterpret_cast<unsigned long long*>(h(f.xsrc)) : nullptr;
        hxdst = xdev ? reinterpret_cast<unsigned long long*>(h(f.xdst)) : nullptr;
        hit = reinterpret_cast<int*>(h(f.it));
        hst = h(f.st);
        hbvv = reinterpret_cast<double*>(h(f.bvv));
        hbv = bv_out ? reinterpret_cast<double*>(h(f.bv)) : nullptr;
        hgap = reinterpret_cast<double*>(h(f.gap));
        hval = reinterpret_cast<double*>(h(f.val));
        if (learned_io) {
            depth = reinterpret_cast<int*>(d(f.depth));
            pc = reinterpret_cast<double*>(d(f.pc));
            topk = reinterpret_cast<int*>(d(f.topk));
            tkx = reinterpret_cast<double*>(d(f.tkx));
            hdepth = reinterpret_cast<int*>(h(f.depth));
            hpc = reinterpret_cast<double*>(h(f.pc));
            htopk = reinterpret_cast<int*>(h(f.topk));
            htkx = reinterpret_cast<double*>(h(f.tkx));
        }
        in_bytes = f.bvv + (size_t)N * 8;
        out_off = xdev ? f.it : f.x;
        out_bytes = f.out_end - out_off;
    }
    gb_status upload() {
        GB_HIP(hipMemcpyAsync(dpack, hpack, in_bytes, hipMemcpyHostToDevice, stream));
        return GB_OK;
    }
    gb_status download() {
        GB_HIP(hipMemcpyAsync(hpack + out_off, dpack + out_off, out_bytes, hipMemcpyDeviceToHost, stream));
        return GB_OK;
    }

    // Buffers grow and never shrink: a batch set is reused by later solves
    // (see BufferPool), so only a larger capacity or n allocates.
    size_t pack_bytes = 0, flags_bytes = 0, feat_elems = 0;

    // Gradients and feature rows (device), pseudocost-free labelling buffers.
    gb_status alloc_features() {
        const size_t xn = (size_t)cap * n;
        if (xn > feat_elems) {
            if (stream) GB_HIP(hipStreamSynchronize(stream));
            for (void* p : {(void*)g, (void*)feat, (void*)score})
                if (p) (void)hipFree(p);
            if (hg) (void)hipHostFree(hg);
            g = nullptr;
            hg = nullptr;
            feat = nullptr;
            score = nullptr;
            feat_elems = 0;
            GB_HIP(hipMalloc(&g, xn * 8));
            GB_HIP(hipHostMalloc(&hg, xn * 8, hipHostMallocDefault));
            GB_HIP(hipMalloc(&feat, xn * kMlpFeatures * 4));
            GB_HIP(hipMalloc(&score, xn * 4));
            feat_elems = xn;
        }
        return GB_OK;
    }

    gb_status alloc(int capacity, int nv, bool learned, int k = 0) {
        cap = capacity;
        n = nv;
        learned_io = learned;
        topk_k = k;
        // each batch set on a hardware queue of its own, so the batches
        // in flight overlap on the device (a solve may then make it share
        // another set's stream instead)
        const bool new_stream = own == nullptr;
        if (new_stream) GB_HIP(create_dedicated_stream(&own));
        stream = own;
        // large enough for every staging format (compact may change per batch)
        const bool c0 = compact, x0 = xdev;
        size_t bytes = 0;
        for (int c = 0; c < 2; ++c)
            for (int d = 0; d < 2; ++d) {
                compact = c != 0;
                xdev = d != 0;
                bytes = std::max(bytes, offsets(cap).end);
            }
        compact = c0;
        xdev = x0;
        if (bytes > pack_bytes) {
            GB_HIP(hipStreamSynchronize(stream));
            if (dpack) (void)hipFree(dpack);
            if (hpack) (void)hipHostFree(hpack);
            dpack = nullptr;
            hpack = nullptr;
            pack_bytes = 0;
            GB_HIP(hipMalloc(&dpack, bytes));
            GB_HIP(hipHostMalloc(&hpack, bytes, hipHostMallocDefault));
            pack_bytes = bytes;
        }
        if ((size_t)n > flags_bytes) {
            if (flags) (void)hipFree(flags);
            flags = nullptr;
            flags_bytes = 0;
            GB_HIP(hipMalloc(&flags, n));
            flags_bytes = (size_t)n;
        }
        if (new_stream) {
            // The runtime creates a stream's hardware queue at its first
            // dispatch (several milliseconds): pay that here, not inside the
            // first search that fills this slot.
            GB_HIP(hipMemsetAsync(flags, 0, (size_t)n, stream));
            GB_HIP(hipStreamSynchronize(stream));
        }
        layout(cap);
        return GB_OK;
    }
    ~DeviceBatch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : {(void*)dpack, (void*)flags, (void*)g, (void*)feat, (void*)score})
            if (p) (void)hipFree(p);
        for (void* p : {(void*)hpack, (void*)hg})
            if (p) (void)hipHostFree(p);
        if (own) (void)hipStreamDestroy(own);
    }
};

// Batch buffer sets (the pipeline slots and the labeller's child batch, with
// their streams and pinned staging) are kept per device and reused by later
// solves: creating streams and pinning host memory per solve cost
// milliseconds now and then.  A solve takes an idle set of its device (or a
// new one) and returns it when it ends; concurrent solves use different sets.
// Sets live until gb_shutdown (or the process exit, which calls it).
// Device iterate store: chunks of NodePool::kChunk slots of n doubles, one
// slot per node-pool slot, kept (like the batch buffers) for later solves.
struct IterateStore {
    std::vector<double*> chunks;
    size_t chunk_bytes = 0;
    // At the start of a solve (no work in flight on the store): chunks of at
    // least `bytes`, covering `slots` slots.  Larger chunks than the store
    // holds replace all of its chunks.
    gb_status reserve(size_t slots, size_t bytes) {
        if (bytes > chunk_bytes) {
            for (double* p : chunks) (void)hipFree(p);
            chunks.clear();
            chunk_bytes = bytes;
        }
        return grow_to(slots);
    }
    // During a solve, while batches in flight read and write the existing
    // chunks: only appends chunks of the current size, never frees one.
    gb_status grow_to(size_t slots) {
        if (chunk_bytes == 0) return fail(GB_ERR_RUNTIME, "iterate store used before reserve");
        while (chunks.size() * (size_t)NodePool::kChunk < slots) {
            double* p = nullptr;
            GB_HIP(hipMalloc(&p, chunk_bytes));
            chunks.push_back(p);
        }
        return GB_OK;
    }
    ~IterateStore() {
        for (double* p : chunks) (void)hipFree(p);
    }
};

constexpr int kMaxSlots = 3;  // pipeline slots (batches in flight) of one search

struct SolveBuffers {
    int device = -1;
    DeviceBatch bufs[kMaxSlots];
    DeviceBatch sb;
    IterateStore xstore;
};

class BufferPool {
  public:
    explicit BufferPool(int device) {
        ensure_shutdown_hook();
        std::lock_guard<std::mutex> lock(mu());
        auto& idle = sets();
        for (size_t i = 0; i < idle.size(); ++i)
            if (idle[i]->device == device) {
                set_ = idle[i];
                idle.erase(idle.begin() + (long)i);
                return;
            }
        set_ = new SolveBuffers();
        set_->device = device;
    }
    ~BufferPool() {
        std::lock_guard<std::mutex> lock(mu());
        sets().push_back(set_);
    }
    SolveBuffers& get() { return *set_; }
    // Frees the idle sets (their destructors wait for their streams).
    static void release_idle() {
        std::vector<SolveBuffers*> idle;
        {
            std::lock_guard<std::mutex> lock(mu());
            idle.swap(sets());
        }
        for (SolveBuffers* b : idle) {
            (void)hipSetDevice(b->device);
            delete b;
        }
    }

  private:
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::vector<SolveBuffers*>& sets() {
        static auto* v = new std::vector<SolveBuffers*>();
        return *v;
    }
    SolveBuffers* set_ = nullptr;
};

gb_status validate_settings(const gb_problem* pr, const gb_objective_desc* o, const gb_lmo_desc* l,
                            const gb_settings* s) {
    if (!pr || !o || !l || !s) return fail(GB_ERR_INVALID_ARGUMENT, "null argument");
    if (pr->n_vars <= 0 || !pr->lb || !pr->ub)
        return fail(GB_ERR_INVALID_ARGUMENT, "problem requires n_vars > 0 and bound vectors");
    // device_resident_solver.cuh:1198-1218
    if (s->gpu_max_iters <= 0 || s->gpu_batch_size <= 0 || s->num_streams <= 0)
        return fail(GB_ERR_INVALID_ARGUMENT, "GPU iterations, batch size and streams must be positive");
    for (int t = 0; t < pr->n_int_vars; ++t)
        if (pr->int_vars[t] < 0 || pr->int_vars[t] >= pr->n_vars)
            return fail(GB_ERR_INVALID_ARGUMENT, "integer variable index out of range");
    if (!std::isfinite(s->gpu_eps) || s->gpu_eps <= 0.0 || !std::isfinite(s->gpu_eps_min)
            || s->gpu_eps_min <= 0.0 || !std::isfinite(s->gpu_eps_decay) || s->gpu_eps_decay <= 0.0
            || s->gpu_eps_decay > 1.0 || !std::isfinite(s->abs_gap_tol) || s->abs_gap_tol < 0.0
            || !std::isfinite(s->rel_gap_tol) || s->rel_gap_tol < 0.0)
        return fail(GB_ERR_INVALID_ARGUMENT, "invalid tolerance configuration");
    if (std::isnan(s->time_limit_seconds) || s->time_limit_seconds <= 0.0)
        return fail(GB_ERR_INVALID_ARGUMENT, "time limit must be positive");
    if (s->gpu_block_threads != 0
            && (s->gpu_block_threads < 64 || s->gpu_block_threads % 64 != 0 || s->gpu_block_threads > 1024))
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "gpu_block_threads must be zero or a positive multiple of 64 up to 1024");
    if (o->n != pr->n_vars)
        return fail(GB_ERR_RUNTIME, "DeviceObjective::n_vars does not match Problem::n_vars");
    // gpuboscia.cuh:86-90
    if (std::isnan(s->natural_lower_bound) || s->natural_lower_bound == kInf)
        return fail(GB_ERR_INVALID_ARGUMENT, "natural lower bound must be finite or negative infinity");
    if (s->branch_rule < GB_BRANCH_POW2 || s->branch_rule > GB_BRANCH_CUSTOM)
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown branch rule");
    if (s->branch_rule == GB_BRANCH_CUSTOM && !s->branch)
        return fail(GB_ERR_INVALID_ARGUMENT, "custom branching requires a branch callback");
    if (s->branch_rule == GB_BRANCH_LEARNED && s->learned_model && o->kind != GB_OBJ_DENSE_QUADRATIC)
        return fail(GB_ERR_INVALID_ARGUMENT, "the learned scorer needs a dense quadratic objective");
    if ((s->branch_rule == GB_BRANCH_POW2 || s->branch_rule == GB_BRANCH_LEARNED)
            && (s->branch_k < 1 || s->branch_k > 16))
        return fail(GB_ERR_INVALID_ARGUMENT, "branch_k must be in [1, 16]");
    if (s->branch_rule == GB_BRANCH_CALLBACK && !s->branch_score)
        return fail(GB_ERR_INVALID_ARGUMENT, "callback branching requires branch_score");
    if (s->form != 0 && s->form != 1 && s->form != 2)
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown solver form");
    if (s->form != 2 && s->precision != GB_PREC_FP64)
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "the plain and batched forms relax in FP64; reduced precision needs form 2");
    if (s->form == 2 && s->precision != GB_PREC_FP64 && s->precision != GB_PREC_FP32
            && s->precision != GB_PREC_BF16)
        return fail(GB_ERR_INVALID_ARGUMENT, "unknown precision");
    if (s->form == 1 && (o->kind != GB_OBJ_DENSE_QUADRATIC || o->n > kMaxFastN
                         || (l->kind != GB_LMO_BOX && !fast_simplex_ok(l))))
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "the batched form supports dense objectives with n <= 512 over a box, or over a simplex "
                    "with an integral right side");
    if (s->form != 0 && l->kind == GB_LMO_SIMPLEX) {
        // the batched and mixed forms' simplex relaxes 0/1 nodes: every
        // variable integer with root bounds 0 or 1 (branching keeps them so)
        bool ok = pr->n_int_vars == pr->n_vars && s->branch_rule != GB_BRANCH_CUSTOM;
        for (int j = 0; j < pr->n_vars && ok; ++j)
            ok = (pr->lb[j] == 0.0 || pr->lb[j] == 1.0) && (pr->ub[j] == 0.0 || pr->ub[j] == 1.0);
        if (!ok)
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "the batched form's simplex needs a 0/1 problem (every variable integer, bounds 0 or 1) "
                        "and no custom branching");
    }
    const bool qap32 = s->form == 2 && o->kind == GB_OBJ_QAP;
    if (qap32 && (l->kind != GB_LMO_BIRKHOFF || l->order != o->order || o->order > kQapMaxOrder
                  || s->device_frontier || s->branch_rule == GB_BRANCH_LEARNED))
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "the QAP's FP32 form relaxes over its Birkhoff polytope (order <= 256), on the host search");
    if (s->form == 2 && !qap32
            && (o->kind != GB_OBJ_DENSE_QUADRATIC || o->n > 512 || (l->kind != GB_LMO_BOX && !fast_simplex_ok(l))))
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "the mixed-precision form supports dense objectives with n <= 512 over a box or a simplex "
                    "with an integral right side, or a QAP over its Birkhoff polytope");
    if (s->form == 2 && l->kind == GB_LMO_SIMPLEX && s->algorithm.algorithm != GB_ALG_DICG)
        return fail(GB_ERR_INVALID_ARGUMENT, "the mixed-precision form runs DICG over a simplex");
    if (s->plugin && (s->form != 0 || s->blocks_per_node > 1 || s->device_frontier))
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "device plug-ins run in the plain form with one workgroup per node (form = 0, SingleBlock)");
    if (s->plugin && s->branch_rule == GB_BRANCH_LEARNED && s->learned_model)
        return fail(GB_ERR_INVALID_ARGUMENT, "the learned scorer needs a dense quadratic objective");
    // device_resident_solver.cuh:1395-1440 (MultiBlock configuration)
    if (s->blocks_per_node < 0) return fail(GB_ERR_INVALID_ARGUMENT, "MultiBlock blocks_per_node must be positive");
    if (s->blocks_per_node > 1) {
        if (s->form != 0)
            return fail(GB_ERR_INVALID_ARGUMENT, "MultiBlock relaxes nodes in the plain form (form = 0)");
        if (o->kind != GB_OBJ_DENSE_QUADRATIC || (l->kind != GB_LMO_BOX && l->kind != GB_LMO_SIMPLEX))
            return fail(GB_ERR_INVALID_ARGUMENT, "MultiBlock relaxes dense quadratics over a box or a simplex");
        const int cap = multiblock_capacity(l->kind);
        if (s->blocks_per_node > kMbMaxK || (cap > 0 && s->blocks_per_node > cap))
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "MultiBlock blocks_per_node exceeds a complete node group of resident workgroups");
    }
    return GB_OK;
}

// validate_root_initial_point (tree.h:115-147)
gb_status validate_root(const gb_problem* pr, const gb_lmo_desc* l, const gb_settings* s) {
    if (!s->root_initial_point || s->root_initial_point_size == 0) return GB_OK;
    const double tol = 1e-9;
    const int n = pr->n_vars;
    if (s->root_initial_point_size != n)
        return fail(GB_ERR_INVALID_ARGUMENT, "invalid root initial point dimensions or tolerance");
    const double* x = s->root_initial_point;
    for (int j = 0; j < n; ++j)
        if (!std::isfinite(x[j]) || std::isnan(pr->lb[j]) || std::isnan(pr->ub[j])
                || pr->lb[j] > pr->ub[j] || x[j] < pr->lb[j] - tol || x[j] > pr->ub[j] + tol)
            return fail(GB_ERR_INVALID_ARGUMENT, "root initial point is nonfinite or violates bounds");
    bool feasible;
    if (s->root_validator) {
        feasible = s->root_validator(s->root_validator_user, x, pr->lb, pr->ub, n, tol) != 0;
    } else if (l->kind == GB_LMO_PLUGIN) {
        if (!s->plugin || !s->plugin->relaxation_feasible)
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "root initial point requires LMO::relaxation_feasible_cpu or "
                        "Settings::root_initial_point_feasible for custom constraints");
        feasible = s->plugin->relaxation_feasible(s->plugin->ctx, x, pr->lb, pr->ub, n, tol) != 0;
    } else {
        feasible = gb_lmo_relaxation_feasible_cpu(l, nullptr, x, pr->lb, pr->ub, n, tol) != 0;
    }
    if (!feasible)
        return fail(GB_ERR_INVALID_ARGUMENT, "root initial point violates the LMO's continuous feasible set");
    return GB_OK;
}

// Strong branching at one node: both children of every fractional candidate
// relaxed in one batched-form launch; appends (features, normalized product
// score) rows.
gb_status strong_branch_labels(const gb_objective_desc* o, const gb_lmo_desc* l, const gb_settings* s,
                               const HostNode& nd, const double* xi, double fval, double gap,
                               NodePool& pool, const Pseudocosts& pc, const std::vector<double>& qdiag,
                               const std::vector<int>& iv, DeviceBatch& sb, Labeler& lab) {
    const int n = pool.n;
    const std::vector<int> cands = fractional_candidates(xi, iv, s->integer_tol);
    if (cands.empty()) return GB_OK;
    const int C = (int)cands.size();
    const int N = 2 * C;
    const double* plb = pool.lb(nd.slot);
    const double* pub = pool.ub(nd.slot);
    const double eps = depth_tolerance(s->gpu_eps, s->gpu_eps_decay, s->gpu_eps_min, nd.depth + 1);
    sb.layout(N);
    for (int c = 0; c < C; ++c)
        for (int d = 0; d < 2; ++d) {
            const int r = 2 * c + d;
            double* lb = sb.hlb + (size_t)r * n;
            double* ub = sb.hub + (size_t)r * n;
            std::copy(plb, plb + n, lb);
            std::copy(pub, pub + n, ub);
            const int v = cands[c];
            if (d) lb[v] = std::ceil(xi[v]);
            else ub[v] = std::floor(xi[v]);
            for (int j = 0; j < n; ++j) sb.hx[(size_t)r * n + j] = std::min(std::max(xi[j], lb[j]), ub[j]);
            sb.heps[r] = eps;
            sb.hit[r] = 0;
            sb.hst[r] = GB_NODE_INITIAL_POINT_SUPPLIED;
            sb.hcut[r] = __builtin_inf();
            sb.hbvv[r] = 1e300;
        }
    gb_status st = sb.upload();
    if (st != GB_OK) return st;
    const gb_fast_options fo{kFastRefresh, 0, 0, nullptr, nullptr};
    st = relax_batch_fast(o, &s->algorithm, l, N, s->gpu_max_iters, sb.x, sb.lb, sb.ub, sb.eps,
                                    sb.it, sb.st, sb.gap, sb.val, sb.bv, sb.bvv, sb.flags,
                                    s->integer_tol, nullptr, &fo, sb.stream);
    if (st != GB_OK) return st;
    // gradient at the parent's iterate for the feature rows
    GB_HIP(hipMemcpyAsync(sb.g, xi, (size_t)n * 8, hipMemcpyHostToDevice, sb.stream));
    const int ge = launch_batched_gradient(o->Qt, o->linear, o->grad_scale, o->linear_scale, n, 1, sb.g,
                                           sb.g + n, sb.stream);
    if (ge != 0) return hip_fail(static_cast<hipError_t>(ge), __FILE__, __LINE__);
    st = sb.download();
    if (st != GB_OK) return st;
    GB_HIP(hipMemcpyAsync(sb.hg, sb.g + n, (size_t)n * 8, hipMemcpyDeviceToHost, sb.stream));
    GB_HIP(hipStreamSynchronize(sb.stream));
    const double big = 1e3 * (1.0 + std::fabs(fval));
    std::vector<double> score(C);
    double smax = 0.0;
    for (int c = 0; c < C; ++c) {
        double gain[2];
        for (int d = 0; d < 2; ++d) {
            const int r = 2 * c + d;
            if (sb.hst[r] == GB_NODE_INFEASIBLE) { gain[d] = big; continue; }
            const double b = certified_fw_lower_bound(nd.bound, sb.hval[r], sb.hgap[r]);
            gain[d] = std::isfinite(b) && std::isfinite(nd.bound) ? std::min(big, std::max(0.0, b - nd.bound))
                                                                   : 0.0;
        }
        score[c] = std::max(gain[0], 1e-6) * std::max(gain[1], 1e-6);
        smax = std::max(smax, score[c]);
    }
    std::vector<float> rows((size_t)C * kMlpFeatures);
    candidate_features(xi, sb.hg, plb, pub, n, nd.depth, gap, fval, pc, qdiag, o->grad_scale, cands,
                       rows.data());
    for (int c = 0; c < C && lab.count < lab.capacity; ++c) {
        std::copy(rows.begin() + (size_t)c * kMlpFeatures, rows.begin() + (size_t)(c + 1) * kMlpFeatures,
                  lab.features + (size_t)lab.count * kMlpFeatures);
        lab.labels[lab.count] = (float)(smax > 0.0 ? score[c] / smax : 0.0);
        ++lab.count;
    }
    ++lab.labeled;
    return GB_OK;
}

gb_status solve_impl(const gb_problem* pr, const gb_objective_desc* o, const gb_lmo_desc* l,
                     const gb_settings* s, gb_result* res, Labeler* lab = nullptr, SolveHooks* hk = nullptr) {
    gb_status st = validate_settings(pr, o, l, s);
    if (st != GB_OK) return st;
    st = validate_descriptors(o, &s->algorithm, l, !(s->form == 0 && s->blocks_per_node > 1), s->plugin);
    if (st != GB_OK) return st;
    if (!(hk && hk->nodes)) {
        st = validate_root(pr, l, s);
        if (st != GB_OK) return st;
    }
    if (s->device_frontier) {
        if (lab || hk)
            return fail(GB_ERR_INVALID_ARGUMENT,
                        "device_frontier runs plain gb_solve searches (no labeller or solver object)");
        return device_search(pr, o, l, s, res);
    }

    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    const int n = pr->n_vars;
    const std::vector<int> iv(pr->int_vars, pr->int_vars + pr->n_int_vars);
    const int B = s->gpu_batch_size;

    const bool learned = s->branch_rule == GB_BRANCH_LEARNED && s->learned_model != nullptr;
    // The batched forms keep several batches in flight (num_streams >= 2):
    // the next batches are chosen, uploaded and launched before the current
    // one is post-processed, so the host's tree work overlaps the GPU's
    // relaxations, and one batch's tail (its slowest nodes) overlaps the
    // others.  Two slots for the batched FP64 form, three for the mixed form
    // (and as many as num_streams allows).  The plain form keeps the
    // reference schedule (one batch at a time).
    //
    // Otherwise (the plain form, or one stream) the slots share one stream,
    // as the reference's single stream slot: when a batch completes, the
    // next full batch is popped and launched before the completed one is
    // post-processed (device_resident_solver.cuh:1585-1596), else the
    // next batch, full or not, follows the post-processing.
    const bool pipelined = s->form != 0 && s->num_streams >= 2;
    const bool shared = !pipelined;
    const int slots = shared ? 2 : std::min(s->form == 2 ? 3 : 2, std::min(s->num_streams, kMaxSlots));
    int device = 0;
    GB_HIP(hipGetDevice(&device));
    // 0/1 problems keep node bounds as bytes (every root bound +0 or 1, and
    // branching rounds fractional values only, so every node bound stays 0
    // or 1); the custom rule and the labeller work on double bounds.
    bool zero_one = s->branch_rule != GB_BRANCH_CUSTOM && !lab;
    {
        // +0.0 or 1.0 exactly (a -0.0 bound would come back as +0.0)
        auto zo = [](double v) { return v == 1.0 || (v == 0.0 && !std::signbit(v)); };
        for (int j = 0; j < n && zero_one; ++j) zero_one = zo(pr->lb[j]) && zo(pr->ub[j]);
        if (hk && hk->nodes)
            for (const InitNode& in : *hk->nodes)
                for (int j = 0; j < n && zero_one; ++j) zero_one = zo(in.lb[j]) && zo(in.ub[j]);
    }
    // Learned searches over 0/1 boxes in the batched forms keep every node's
    // iterate on the device: the learned rule takes its branching variables
    // from the device, so the host never needs x (GPUBOSCIA_HOST_ITERATES=1
    // stages iterates through the host instead; the search is the same).
    const char* host_iterates = std::getenv("GPUBOSCIA_HOST_ITERATES");
    const bool xdev = learned && zero_one && s->form != 0 && !(host_iterates && host_iterates[0] == '1');
    BufferPool pooled(device);
    DeviceBatch* bufs = pooled.get().bufs;
    IterateStore& xstore = pooled.get().xstore;
    const size_t xchunk_bytes = (size_t)NodePool::kChunk * n * 8;
    for (int k = 0; k < slots; ++k) {
        bufs[k].compact = zero_one;
        bufs[k].xdev = xdev;
        st = bufs[k].alloc(B, n, learned, s->branch_k);
        if (st != GB_OK) return st;
        if (learned || lab) {
            st = bufs[k].alloc_features();
            if (st != GB_OK) return st;
        }
    }
    if (shared)
        for (int k = 1; k < slots; ++k) bufs[k].stream = bufs[0].stream;
    // Child batch of the strong-branching labeller (2 children per candidate).
    DeviceBatch& sb = pooled.get().sb;
    if (lab) {
        st = sb.alloc(2 * n, n, false);
        if (st != GB_OK) return st;
        st = sb.alloc_features();
        if (st != GB_OK) return st;
    }
    Pseudocosts pc;
    pc.init(n);
    // Q diagonal on the host for the labeller's features (feature estimate
    // of branch gains); the device features read it from Qt directly.
    std::vector<double> qdiag;
    if (lab && o->kind == GB_OBJ_DENSE_QUADRATIC) {
        qdiag.resize(n);
        GB_HIP(hipMemcpy2D(qdiag.data(), sizeof(double), o->Qt, (size_t)(n + 1) * sizeof(double),
                           sizeof(double), n, hipMemcpyDeviceToHost));
    }
    std::vector<uint8_t> hflags(n, 0);
    for (int i : iv) hflags[i] = 1;
    for (int k = 0; k < slots; ++k) GB_HIP(hipMemcpy(bufs[k].flags, hflags.data(), n, hipMemcpyHostToDevice));
    // Mixed-precision form: Q split into bf16 planes once per solve (one
    // plane for GB_PREC_BF16, three otherwise).
    std::unique_ptr<gb_mixed_q, gb_status (*)(gb_mixed_q*)> mq(nullptr, gb_mixed_destroy);
    // The QAP's FP32 form: FP32 copies of A, B and A'A, B B' once per solve.
    std::unique_ptr<gb_qap32, gb_status (*)(gb_qap32*)> qq(nullptr, gb_qap32_destroy);
    const bool qap32 = s->form == 2 && o->kind == GB_OBJ_QAP;
    const double qstop = std::isfinite(s->natural_lower_bound) ? s->natural_lower_bound + s->abs_gap_tol
                                                               : -std::numeric_limits<double>::infinity();
    if (qap32) {
        gb_qap32* h = nullptr;
        st = gb_qap32_prepare(o, bufs[0].stream, &h);
        if (st != GB_OK) return st;
        qq.reset(h);
        GB_HIP(hipStreamSynchronize(bufs[0].stream));
    } else if (s->form == 2) {
        gb_mixed_q* h = nullptr;
        st = gb_mixed_prepare(o, s->precision == GB_PREC_BF16 ? 1 : 3, bufs[0].stream, &h);
        if (st != GB_OK) return st;
        mq.reset(h);
        GB_HIP(hipStreamSynchronize(bufs[0].stream));  // every slot's stream uses the planes
    }
    gb_mixed_options tune{0, 0, 0, 1e-6f, 0, nullptr, nullptr, nullptr};
    bool tuned = false;
    // Every exit drains every slot's stream before the mixed planes and the
    // batch buffers are released (a second batch may still be in flight).
    struct DrainOnExit {
        DeviceBatch* b;
        int count;
        ~DrainOnExit() {
            for (int k = 0; k < count; ++k)
                if (b[k].stream) (void)hipStreamSynchronize(b[k].stream);
        }
    } drain{bufs, slots};
    if (lab) GB_HIP(hipMemcpy(sb.flags, hflags.data(), n, hipMemcpyHostToDevice));

    Frontier q;
    NodePool pool;
    pool.n = n;
    pool.compact = zero_one;
    if (xdev) {
        st = xstore.reserve(1, xchunk_bytes);
        if (st != GB_OK) return st;
        pool.dchunks = &xstore.chunks;
    }
    // The initial nodes: the root of the problem (gpuboscia.cuh:86-96), or
    // the nodes pushed on a solver object.
    auto add_initial = [&](int id, int parent, int depth, const double* lb, const double* ub, const double* rx,
                           int riters, double bound) -> gb_status {
        HostNode nd;
        nd.id = id;
        nd.parent = parent;
        nd.depth = depth;
        nd.slot = pool.alloc();
        if (pool.compact) {
            for (int j = 0; j < n; ++j) {
                pool.lb8(nd.slot)[j] = lb[j] == 1.0 ? 1 : 0;
                pool.ub8(nd.slot)[j] = ub[j] == 1.0 ? 1 : 0;
            }
        } else {
            std::copy(lb, lb + n, pool.lb(nd.slot));
            std::copy(ub, ub + n, pool.ub(nd.slot));
        }
        if (rx) {
            if (xdev)
                GB_HIP(hipMemcpy(reinterpret_cast<double*>(pool.drx(nd.slot)), rx, (size_t)n * 8,
                                 hipMemcpyHostToDevice));
            else
                std::copy(rx, rx + n, pool.rx(nd.slot));
            nd.has_rx = true;
        }
        nd.riters = riters;
        nd.bound = bound;
        q.push(nd);
        return GB_OK;
    };
    if (hk && hk->nodes) {
        for (const InitNode& in : *hk->nodes) {
            st = add_initial(in.id, in.parent, in.depth, in.lb.data(), in.ub.data(),
                             in.rx.empty() ? nullptr : in.rx.data(), in.riters, in.bound);
            if (st != GB_OK) return st;
        }
    } else {
        const bool has_rx = s->root_initial_point && s->root_initial_point_size == n;
        st = add_initial(0, -1, 0, pr->lb, pr->ub, has_rx ? s->root_initial_point : nullptr, 0,
                         s->natural_lower_bound);
        if (st != GB_OK) return st;
    }
    double inc = kInf;
    std::vector<double> inc_x;
    int next_id = 1, done_nodes = 0;
    long long fw_iters = 0;
    double final_bound = -kInf, relax_s = 0.0;
    bool tol_hit = false, time_hit = false, finished = false;
    // Bounds of the nodes of each slot's batch not yet post-processed: nodes
    // leave in batch order, so the smallest is the suffix minimum at the
    // slot's first open position (pos == size: nothing in flight).
    std::vector<double> open_min[kMaxSlots];
    size_t open_pos[kMaxSlots] = {};
    std::vector<HostNode> batches[kMaxSlots];
    // device iterates: parents' slots read by each slot's gather, released
    // once that batch has finished
    std::vector<int> gathered[kMaxSlots];
    auto in_flight = [&](int k) { return open_pos[k] < open_min[k].size(); };

    auto live_bound = [&]() {
        double b = inc;
        if (!q.empty()) b = std::min(b, q.top_bound());
        for (int k = 0; k < slots; ++k)
            if (in_flight(k)) b = std::min(b, open_min[k][open_pos[k]]);
        return b;
    };
    auto finish_check = [&](bool batch_remaining) {
        if (q.empty() && !batch_remaining) { final_bound = inc; finished = true; return; }
        const double b = live_bound();
        if (gap_within_tolerance(b, inc, s->abs_gap_tol, s->rel_gap_tol)) {
            final_bound = b;
            tol_hit = true;
            finished = true;
        }
    };
    auto progress = [&](bool new_inc, int open_in_batch) {
        if (!s->progress) return;
        s->progress(s->progress_user, done_nodes, (int)q.size() + open_in_batch, live_bound(), inc,
                    new_inc ? 1 : 0);
    };

    // Select the next batch from the queue into `slot`, upload it and queue
    // its kernels and the result copy on the slot's stream (an empty queue
    // leaves the slot idle).
    // Scheduling observer (device_resident_solver.cuh:1027-1038): events of
    // the stream slot a batch runs on, with that slot's dispatch count.
    unsigned long long generation[kMaxSlots] = {};  // dispatches per stream slot
    unsigned long long batch_gen[kMaxSlots] = {};   // the generation of each buffer set's batch
    auto emit = [&](int type, int slot, int count) {
        if (!hk || !hk->sched) return;
        const gb_scheduling_event ev{type, shared ? 0 : slot, batch_gen[slot], count};
        hk->sched(hk->sched_user, &ev);
    };
    auto launch_batch = [&](int slot) -> gb_status {
        gb_status st = GB_OK;
        std::vector<HostNode>& batch = batches[slot];
        DeviceBatch& db = bufs[slot];
        batch.clear();
        open_min[slot].clear();
        open_pos[slot] = 0;
        while (!q.empty() && (int)batch.size() < B) {
            HostNode nd = q.top();
            q.pop();
            if (nd.bound >= inc) {
                pool.release(nd.slot);
                if (nd.xslot >= 0) pool.release(nd.xslot);
                continue;
            }
            batch.push_back(nd);
        }
        if (batch.empty()) return GB_OK;
        const int N = (int)batch.size();
        // bytes on the bus when the batch is large enough to pay for the
        // widening launch (about 16 K bound entries); doubles otherwise
        db.compact = pool.compact && (size_t)N * n >= 16384;
        db.bv_out = (size_t)N * n < 16384;
        db.layout(N);
        if (db.xdev) {
            // store slots for every pool slot handed out so far (other
            // slots' batches may be using the store)
            st = xstore.grow_to((size_t)pool.next);
            if (st != GB_OK) return st;
        }
        // Staging copies on the host threads (the pool is only read here).
        HostThreads::get().parallel_for(N, [&](int i) {
            const HostNode& nd = batch[i];
            double* hx = db.xdev ? nullptr : db.hx + (size_t)i * n;
            const bool warm = nd.has_rx || nd.xslot >= 0;
            if (db.xdev) {
                // start point: own iterate as is, the parent's clipped (bit 0), or lb
                db.hxsrc[i] = nd.has_rx ? pool.drx(nd.slot) : nd.xslot >= 0 ? (pool.drx(nd.xslot) | 1ull) : 0ull;
                db.hxdst[i] = pool.drx(nd.slot);
            }
            if (pool.compact) {
                const uint8_t* l8 = pool.lb8(nd.slot);
                const uint8_t* u8 = pool.ub8(nd.slot);
                if (db.compact) {
                    // bytes travel; the device widens them (0/1 values, exact)
                    std::copy(l8, l8 + n, db.hlb8 + (size_t)i * n);
                    std::copy(u8, u8 + n, db.hlb8 + (size_t)(N + i) * n);
                } else {
                    double* hl = db.hlb + (size_t)i * n;
                    double* hu = db.hub + (size_t)i * n;
                    for (int j = 0; j < n; ++j) {
                        hl[j] = l8[j];
                        hu[j] = u8[j];
                    }
                }
                if (db.xdev) {
                    // gathered on the device
                } else if (nd.has_rx) {
                    std::copy(pool.rx(nd.slot), pool.rx(nd.slot) + n, hx);
                } else if (nd.xslot >= 0) {
                    const double* px = pool.rx(nd.xslot);
                    for (int j = 0; j < n; ++j)
                        hx[j] = std::min(std::max(px[j], (double)l8[j]), (double)u8[j]);
                } else {
                    for (int j = 0; j < n; ++j) hx[j] = l8[j];
                }
            } else {
                std::copy(pool.lb(nd.slot), pool.lb(nd.slot) + n, db.hlb + (size_t)i * n);
                std::copy(pool.ub(nd.slot), pool.ub(nd.slot) + n, db.hub + (size_t)i * n);
                const double* nlb = db.hlb + (size_t)i * n;
                const double* nub = db.hub + (size_t)i * n;
                if (nd.has_rx) {
                    std::copy(pool.rx(nd.slot), pool.rx(nd.slot) + n, hx);
                } else if (nd.xslot >= 0) {
                    // Batched forms: the parent's iterate clipped to this box (any
                    // box point is feasible).
                    const double* px = pool.rx(nd.xslot);
                    for (int j = 0; j < n; ++j) hx[j] = std::min(std::max(px[j], nlb[j]), nub[j]);
                } else {
                    std::copy(nlb, nlb + n, hx);
                }
            }
            db.heps[i] = depth_tolerance(s->gpu_eps, s->gpu_eps_decay, s->gpu_eps_min, nd.depth);
            db.hit[i] = nd.riters;
            db.hst[i] = warm ? GB_NODE_INITIAL_POINT_SUPPLIED : 0;
            db.hcut[i] = inc;
            db.hbvv[i] = 1e300;
            if (learned) db.hdepth[i] = nd.depth;
        });
        for (HostNode& nd : batch)
            if (nd.xslot >= 0) {
                if (db.xdev) gathered[slot].push_back(nd.xslot);  // the gather below reads it
                else pool.release(nd.xslot);
                nd.xslot = -1;
            }
        open_min[slot].resize(N);
        for (int i = N - 1; i >= 0; --i)
            open_min[slot][i] = i + 1 < N ? std::min(batch[i].bound, open_min[slot][i + 1]) : batch[i].bound;
        if (learned) {
            for (int j = 0; j < n; ++j) {
                db.hpc[j] = pc.up_cnt[j] ? pc.up_sum[j] / pc.up_cnt[j] : -1.0;
                db.hpc[n + j] = pc.dn_cnt[j] ? pc.dn_sum[j] / pc.dn_cnt[j] : -1.0;
            }
        }
        const auto r0 = Clock::now();
        // One copy in (bounds, tolerances, cutoffs, warm starts, counters,
        // learned inputs), one copy out after the kernels.
        st = db.upload();
        if (st != GB_OK) return st;
        if (db.compact) {
            const int ee = launch_expand_bounds((size_t)N * n, db.lb8, db.lb8 + (size_t)N * n, db.lb, db.ub,
                                               db.stream);
            if (ee != 0) return hip_fail(static_cast<hipError_t>(ee), __FILE__, __LINE__);
        }
        if (db.xdev) {
            const int ge = launch_gather_starts(n, N, db.xsrc, db.lb, db.ub, db.x, db.stream);
            if (ge != 0) return hip_fail(static_cast<hipError_t>(ge), __FILE__, __LINE__);
        }
        if (s->form == 1) {
            const gb_fast_options fo{kFastRefresh, 0, 0, nullptr, learned ? db.g : nullptr};
            st = relax_batch_fast(o, &s->algorithm, l, N, s->gpu_max_iters, db.x, db.lb, db.ub,
                                  db.eps, db.it, db.st, db.gap, db.val, db.bv, db.bvv, db.flags,
                                  s->integer_tol, db.cut, &fo, db.stream);
        } else if (qap32) {
            st = gb_relax_batch_qap32(qq.get(), l, N, s->gpu_max_iters, db.x, db.lb, db.ub, db.eps, db.it, db.st,
                                      db.gap, db.val, db.bv, db.bvv, db.flags, s->integer_tol, qstop, 0,
                                      db.stream);
        } else if (s->form == 2) {
            if (s->autotune && !tuned && N >= 64) {
                st = gb_mixed_autotune(mq.get(), o, &s->algorithm, l, N, s->gpu_max_iters, db.x, db.lb, db.ub,
                                       db.eps, db.it, db.st, db.flags, s->integer_tol, db.cut, &tune, nullptr,
                                       db.stream);
                if (st != GB_OK) return st;
                tuned = true;
            }
            gb_mixed_options mo = tune;
            mo.gradient = learned ? db.g : nullptr;
            st = gb_relax_batch_mixed(mq.get(), o, &s->algorithm, l, N, s->gpu_max_iters, db.x, db.lb,
                                      db.ub, db.eps, db.it, db.st, db.gap, db.val, db.bv, db.bvv,
                                      db.flags, s->integer_tol, db.cut, &mo, db.stream);
        } else if (s->plugin) {
            // the caller's own objective / LMO / heuristic types in the same kernel
            st = relax_batch_plugin(o, &s->algorithm, l, s->plugin, N, s->gpu_max_iters, db.x, db.lb, db.ub,
                                    db.eps, db.it, db.st, db.gap, db.val, db.bv, db.bvv, db.flags,
                                    s->integer_tol, s->gpu_block_threads, db.stream);
        } else if (s->blocks_per_node > 1) {
            // MultiBlock (FWExecutionMode::MultiBlock): the same relaxation,
            // bit for bit, with blocks_per_node workgroups per node
            st = relax_batch_multiblock(o, &s->algorithm, l, N, s->gpu_max_iters, db.x, db.lb, db.ub,
                                        db.eps, db.it, db.st, db.gap, db.val, db.bv, db.bvv, db.flags,
                                        s->integer_tol, s->gpu_block_threads, s->blocks_per_node, db.stream);
        } else {
            st = relax_batch_fp64(o, &s->algorithm, l, N, s->gpu_max_iters, db.x, db.lb, db.ub,
                                  db.eps, db.it, db.st, db.gap, db.val, db.bv, db.bvv, db.flags,
                                  s->integer_tol, s->gpu_block_threads, db.stream);
        }
        if (st != GB_OK) return st;
        if (learned) {
            // Scores of every (node, variable) on the device, returned with
            // the batch: gradient (from the relaxation kernels, or one
            // product launch for the plain form), feature rows, MLP.
            if (s->form == 0) {
                const int ge = launch_batched_gradient(o->Qt, o->linear, o->grad_scale, o->linear_scale,
                                                       n, N, db.x, db.g, db.stream);
                if (ge != 0) return hip_fail(static_cast<hipError_t>(ge), __FILE__, __LINE__);
            }
            const int fe = launch_candidate_features(n, N, db.x, db.g, db.lb, db.ub, db.depth, db.gap, db.val,
                                                     db.pc, db.pc + n, o->Qt, (size_t)n + 1, o->grad_scale, db.flags,
                                                     s->integer_tol, db.feat, db.stream);
            if (fe != 0) return hip_fail(static_cast<hipError_t>(fe), __FILE__, __LINE__);
            st = gb_brancher_predict(s->learned_model, db.feat, N * n, db.score, db.stream);
            if (st != GB_OK) return st;
            const int te = launch_topk_candidates(n, N, db.x, db.flags, s->integer_tol, db.score, db.st,
                                                  s->branch_k, db.topk, db.tkx, db.stream);
            if (te != 0) return hip_fail(static_cast<hipError_t>(te), __FILE__, __LINE__);
        }
        if (db.xdev) {
            const int se = launch_scatter_iterates(n, N, db.xdst, db.x, db.stream);
            if (se != 0) return hip_fail(static_cast<hipError_t>(se), __FILE__, __LINE__);
        }
        st = db.download();
        if (st != GB_OK) return st;
        relax_s += std::chrono::duration<double>(Clock::now() - r0).count();
        batch_gen[slot] = ++generation[shared ? 0 : slot];
        emit(GB_SCHED_DISPATCHED, slot, N);
        return GB_OK;
    };

    struct ChildTask {
        int child, parent, row, mask, k, var0;
    };
    struct XTask {
        int slot, row;
    };
    std::vector<ChildTask> child_tasks;
    std::vector<XTask> x_tasks;
    std::vector<int> deferred, task_vars, scan_vars, vars;  // vars: one node's branching variables
    std::vector<double> xrow;                               // one node's x fetched from the device
    std::vector<HostNode> kids;                             // one node's children, pushed as a family
    int head = 0;
    while (!finished) {
        const double el = std::chrono::duration<double>(Clock::now() - t0).count();
        if (el >= s->time_limit_seconds || (s->max_nodes > 0 && done_nodes >= s->max_nodes)) {
            final_bound = live_bound();
            time_hit = true;
            break;
        }
        if (!in_flight(head)) {
            st = launch_batch(head);
            if (st != GB_OK) return st;
            if (!in_flight(head)) { final_bound = inc; finished = true; break; }
        }
        // Separate streams: the other slots in launch order after head, idle
        // ones take the next batches (full or not) before head completes.
        if (!shared)
            for (int d = 1; d < slots; ++d) {
                const int k = (head + d) % slots;
                if (!in_flight(k)) {
                    st = launch_batch(k);
                    if (st != GB_OK) return st;
                }
            }
        DeviceBatch& db = bufs[head];
        std::vector<HostNode>& batch = batches[head];
        const int N = (int)batch.size();
        {
            const auto w0 = Clock::now();
            GB_HIP(hipStreamSynchronize(db.stream));
            relax_s += std::chrono::duration<double>(Clock::now() - w0).count();
        }
        // status 0xFF: a MultiBlock group whose workgroups were not resident
        // together gave up at its barrier, so the node was not relaxed
        for (int i = 0; i < N && s->form == 0 && s->blocks_per_node > 1; ++i)
            if (db.hst[i] > GB_NODE_INTEGRAL_CONVERGED)
                return fail(GB_ERR_DEVICE, "a node's relaxation did not complete (status "
                                               + std::to_string(db.hst[i])
                                               + "): its MultiBlock workgroups were not resident together");
        // status 0xFF from the batched form's simplex: a node outside its 0/1 scope
        for (int i = 0; i < N && s->form != 0 && l->kind == GB_LMO_SIMPLEX; ++i)
            if (db.hst[i] > GB_NODE_INTEGRAL_CONVERGED)
                return fail(GB_ERR_DEVICE, "the batched form's simplex met a node that is not 0/1");
        emit(GB_SCHED_COMPLETION_DETACHED, head, N);
        for (int sl : gathered[head]) pool.release(sl);
        gathered[head].clear();
        // One stream: its next batch, if a full one is waiting, is launched
        // before this one is post-processed.
        if (shared) {
            const int k = (head + 1) % slots;
            if (!in_flight(k) && (int)q.size() >= B) {
                st = launch_batch(k);
                if (st != GB_OK) return st;
            }
        }
        // Nodes of the other slots count as open while head's are processed.
        int queued_elsewhere = 0;
        for (int d = 1; d < slots; ++d) {
            const int k = (head + d) % slots;
            if (in_flight(k)) queued_elsewhere += (int)batches[k].size();
        }
        emit(GB_SCHED_POSTPROCESS_STARTED, head, N);

        for (int i = 0; i < N; ++i) fw_iters += db.hit[i] - batch[i].riters;
        // Branching variables of the rules that depend only on x, on the host
        // threads; child bounds and iterates are copied after the sequential
        // pass (parent slots are released only then).
        const bool prescan = s->branch_rule == GB_BRANCH_POW2 || s->branch_rule == GB_BRANCH_MOST_FRACTIONAL ||
                             (s->branch_rule == GB_BRANCH_LEARNED && !learned);
        const int bk = s->branch_rule == GB_BRANCH_POW2 ? s->branch_k : 1;
        if (prescan) {
            scan_vars.assign((size_t)N * bk, -1);
            HostThreads::get().parallel_for(N, [&](int i) {
                if (db.hst[i] >= GB_NODE_INTEGRAL_UNCONVERGED) return;
                // pruned below whatever incumbent the pass finds (it only improves)
                if (!(certified_fw_lower_bound(batch[i].bound, db.hval[i], db.hgap[i]) < inc)) return;
                const double* xr = db.hx + (size_t)i * n;
                int* out = scan_vars.data() + (size_t)i * bk;
                if (s->branch_rule == GB_BRANCH_POW2) {
                    thread_local std::vector<std::pair<int, double>> cand;
                    top_k_infeasible_into(xr, iv, s->branch_k, s->integer_tol, out, cand);
                } else {
                    out[0] = most_fractional(xr, iv, s->integer_tol);
                }
            });
        }
        child_tasks.clear();
        x_tasks.clear();
        deferred.clear();
        task_vars.clear();
        for (int i = 0; i < N && !finished; ++i) {
            HostNode nd = batch[i];
            open_pos[head] = (size_t)i + 1;
            const int remaining = N - i - 1 + queued_elsewhere;
            if (db.hst[i] == GB_NODE_INFEASIBLE) {
                pool.release(nd.slot);
                ++done_nodes;
                progress(false, remaining);
                finish_check(remaining > 0);
                continue;
            }
            const double node_eps = db.heps[i];
            const double g = db.hgap[i];
            if (hk) {  // device_resident_solver.cuh:1931-1937
                hk->last_iters = db.hit[i];
                hk->last_gap = g;
                hk->last_conv = std::isfinite(g) && g >= 0.0 && g < node_eps;
            }
            nd.bound = certified_fw_lower_bound(nd.bound, db.hval[i], g);
            if (nd.bvar >= 0 && nd.riters == 0 && nd.bfrac > 0.0 && std::isfinite(nd.bound)
                    && std::isfinite(nd.pbound))
                pc.record(nd.bvar, nd.bdir, std::max(0.0, nd.bound - nd.pbound) / nd.bfrac);
            bool new_inc = false;
            const double bvv = db.hbvv[i];
            if (bvv < 1e299 && std::isfinite(bvv) && bvv < inc) {
                inc = bvv;
                if (db.hbv) {
                    inc_x.assign(db.hbv + (size_t)i * n, db.hbv + (size_t)(i + 1) * n);
                } else {
                    // the batch's stream is idle here (its results were waited for)
                    inc_x.resize(n);
                    GB_HIP(hipMemcpyAsync(inc_x.data(), db.bv + (size_t)i * n, (size_t)n * 8,
                                          hipMemcpyDeviceToHost, db.stream));
                    GB_HIP(hipStreamSynchronize(db.stream));
                }
                new_inc = true;
            }
            const bool cert = std::isfinite(g) && g >= 0.0 && g < node_eps;
            if (nd.bound >= inc) {
                pool.release(nd.slot);
                ++done_nodes;
                progress(new_inc, remaining);
                finish_check(remaining > 0);
                continue;
            }
            // null with device iterates (the scatter kept x in the node's store slot)
            const double* xi = db.hx ? db.hx + (size_t)i * n : nullptr;
            if (db.hst[i] >= GB_NODE_INTEGRAL_UNCONVERGED && !cert) {
                // device_resident_solver.cuh:1970-1980: keep the node live.
                if (!db.xdev) x_tasks.push_back({nd.slot, i});
                nd.has_rx = true;
                nd.riters = db.hit[i];
                q.push(nd);
                progress(new_inc, remaining);
                continue;
            }
            if (db.hst[i] < GB_NODE_INTEGRAL_UNCONVERGED && s->branch_rule == GB_BRANCH_CUSTOM) {
                // Host BranchStrategy (device_resident_solver.cuh:1986-1995):
                // the node as launched, with its new bound; 8 ids reserved.
                const gb_node view{nd.id, nd.parent, nd.depth, pool.lb(nd.slot), pool.ub(nd.slot),
                                   nd.has_rx ? pool.rx(nd.slot) : nullptr, nd.riters, nd.bound};
                gb_children kids_out;
                kids_out.n = n;
                int nid = next_id;
                next_id += 8;
                if (s->branch(s->branch_user, &view, &nid, xi, n, iv.data(), (int)iv.size(), &kids_out) != 0)
                    return fail(GB_ERR_RUNTIME, "the branch callback failed");
                if (nid > next_id) next_id = nid;
                for (const gb_children::Child& ch : kids_out.kids) {
                    HostNode c;
                    c.id = ch.id;
                    c.parent = ch.parent;
                    c.depth = ch.depth;
                    c.bound = ch.bound;
                    if (!(c.bound < inc)) continue;
                    c.slot = pool.alloc();
                    std::copy(ch.lb.begin(), ch.lb.end(), pool.lb(c.slot));
                    std::copy(ch.ub.begin(), ch.ub.end(), pool.ub(c.slot));
                    q.push(c);
                }
            } else if (db.hst[i] < GB_NODE_INTEGRAL_UNCONVERGED) {
                vars.clear();
                const double* xr = xi;        // x of the node, where the host has it
                const double* tkx = nullptr;  // x at the chosen candidates (learned rule)
                if (prescan) {
                    const int* sv = scan_vars.data() + (size_t)i * bk;
                    for (int c = 0; c < bk && sv[c] >= 0; ++c) vars.push_back(sv[c]);
                } else if (s->branch_rule == GB_BRANCH_CALLBACK) {
                    std::vector<int> cand;
                    for (int v : iv) {
                        const double frac = xi[v] - std::floor(xi[v]);
                        if (std::min(frac, 1.0 - frac) > s->integer_tol) cand.push_back(v);
                    }
                    if (!cand.empty()) {
                        std::vector<double> score(cand.size(), 0.0);
                        s->branch_score(s->branch_score_user, xi, n, cand.data(), (int)cand.size(),
                                        nd.depth, score.data());
                        size_t bi = 0;
                        for (size_t c = 1; c < cand.size(); ++c)
                            if (score[c] > score[bi]) bi = c;
                        vars.push_back(cand[bi]);
                    }
                } else if (learned && db.htopk[(size_t)i * s->branch_k] >= 0) {
                    // the branch_k candidates with the highest predicted
                    // strong-branching score (ties to the lower index, chosen
                    // on the device), split 2^k ways like branch_pow2
                    const int* tk = db.htopk + (size_t)i * s->branch_k;
                    for (int c = 0; c < s->branch_k && tk[c] >= 0; ++c) vars.push_back(tk[c]);
                    tkx = db.htkx + (size_t)i * s->branch_k;
                } else {
                    // the learned rule where no candidate was scored
                    if (!xr) {
                        // device iterates: this node's row (the stream is idle)
                        xrow.resize(n);
                        GB_HIP(hipMemcpyAsync(xrow.data(), db.x + (size_t)i * n, (size_t)n * 8,
                                              hipMemcpyDeviceToHost, db.stream));
                        GB_HIP(hipStreamSynchronize(db.stream));
                        xr = xrow.data();
                    }
                    const int v = most_fractional(xr, iv, s->integer_tol);
                    if (v >= 0) vars.push_back(v);
                }
                if (lab && lab->labeled < lab->max_nodes) {
                    st = strong_branch_labels(o, l, s, nd, xi, db.hval[i], db.hgap[i], pool, pc, qdiag,
                                              iv, sb, *lab);
                    if (st != GB_OK) return st;
                }
                const int base = next_id;
                next_id += 8;  // device_resident_solver.cuh:1989
                int nid = base;
                const int k = (int)vars.size();
                const int var0 = (int)task_vars.size();
                kids.clear();
                task_vars.insert(task_vars.end(), vars.begin(), vars.end());
                bool parent_x_kept = false;
                for (int mask = 0; k > 0 && mask < (1 << k); ++mask) {
                    HostNode c;
                    c.id = nid++;
                    c.parent = nd.id;
                    c.depth = nd.depth + 1;
                    c.bound = nd.bound;
                    if (!(c.bound < inc)) continue;
                    c.slot = pool.alloc();
                    child_tasks.push_back({c.slot, nd.slot, i, mask, k, var0});
                    if (s->form != 0 && l->kind != GB_LMO_SIMPLEX) {
                        // Batched forms: the child starts from the parent's
                        // iterate, kept once in the parent's slot and clipped to
                        // the child's box when the child is staged (a box point
                        // is feasible; over a simplex the clipped point may not
                        // be, so those children start at the LMO vertex).
                        if (!parent_x_kept) {
                            if (!db.xdev) x_tasks.push_back({nd.slot, i});
                            parent_x_kept = true;
                        }
                        pool.retain(nd.slot);
                        c.xslot = nd.slot;
                    }
                    if (k == 1) {
                        const int vi = vars[0];
                        c.bvar = vi;
                        c.bdir = mask & 1;
                        const double xv = tkx ? tkx[0] : xr[vi];
                        c.bfrac = c.bdir ? std::ceil(xv) - xv : xv - std::floor(xv);
                        c.pbound = nd.bound;
                    }
                    kids.push_back(c);
                }
                q.push_family(kids.data(), (int)kids.size());
            }
            deferred.push_back(nd.slot);
            ++done_nodes;
            progress(new_inc, remaining);
            finish_check(remaining > 0);
        }
        // Child bounds (the parent's with the branched variables rounded) and
        // kept iterates, on the host threads; then the parents' slots go back.
        HostThreads::get().parallel_for((int)child_tasks.size(), [&](int t) {
            const ChildTask& ct = child_tasks[t];
            if (pool.compact) {
                // branched values are fractional, so ceil is 1 and floor 0
                uint8_t* cl8 = pool.lb8(ct.child);
                std::copy(pool.lb8(ct.parent), pool.lb8(ct.parent) + 2 * n, cl8);  // lb and ub
                for (int b = 0; b < ct.k; ++b) {
                    const int vi = task_vars[ct.var0 + b];
                    if (ct.mask & (1 << b)) cl8[vi] = 1;
                    else cl8[n + vi] = 0;
                }
                return;
            }
            const double* xr = db.hx + (size_t)ct.row * n;
            double* clb = pool.lb(ct.child);
            double* cub = pool.ub(ct.child);
            std::copy(pool.lb(ct.parent), pool.lb(ct.parent) + 2 * n, clb);  // lb and ub
            for (int b = 0; b < ct.k; ++b) {
                const int vi = task_vars[ct.var0 + b];
                if (ct.mask & (1 << b)) clb[vi] = std::ceil(xr[vi]);
                else cub[vi] = std::floor(xr[vi]);
            }
        });
        HostThreads::get().parallel_for((int)x_tasks.size(), [&](int t) {
            const double* xr = db.hx + (size_t)x_tasks[t].row * n;
            std::copy(xr, xr + n, pool.rx(x_tasks[t].slot));
        });
        for (int sl : deferred) pool.release(sl);
        // next: the oldest batch in flight (slots are launched cyclically)
        int next = -1;
        for (int d = 1; d < slots && next < 0; ++d)
            if (in_flight((head + d) % slots)) next = (head + d) % slots;
        if (!finished && q.empty() && next < 0) { final_bound = inc; finished = true; }
        if (next >= 0) head = next;
    }
    if (final_bound == -kInf) final_bound = q.empty() ? inc : std::min(inc, q.top_bound());
    res->objective = inc;
    res->dual_bound = final_bound;
    res->nodes_processed = done_nodes;
    res->elapsed_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
    res->reached_tolerance = tol_hit ? 1 : 0;
    res->reached_time_limit = time_hit ? 1 : 0;
    res->fw_iterations = fw_iters;
    res->relax_seconds = relax_s;
    if (res->x)
        for (int j = 0; j < n; ++j) res->x[j] = inc_x.empty() ? std::nan("") : inc_x[j];
    return GB_OK;
}

}  // namespace

double depth_tolerance(double root, double decay, double minimum, int depth) {
    if (depth <= 0) return std::max(root, minimum);
    return std::max(minimum, root * std::pow(decay, depth));
}

void release_buffer_sets() { BufferPool::release_idle(); }

}  // namespace gb

extern "C" void gb_default_settings(gb_settings* s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->gpu_batch_size = 1024;
    s->num_streams = 8;
    s->gpu_block_threads = 0;
    s->gpu_max_iters = 20000;
    s->gpu_eps = 1e-3;
    s->gpu_eps_decay = 0.8;
    s->gpu_eps_min = 1e-6;
    s->integer_tol = 1e-6;
    s->abs_gap_tol = 1e-6;
    s->rel_gap_tol = 1e-2;
    s->algorithm.algorithm = GB_ALG_DICG;
    s->algorithm.line_search = GB_LS_SECANT;
    s->algorithm.secant_iters = 4;
    s->algorithm.secant_tol = 1e-10;
    s->algorithm.secant_denom_eps = 1e-14;
    s->algorithm.heuristic = GB_HEUR_ROUND;
    s->algorithm.heuristic_offset = 0.0;
    s->branch_rule = GB_BRANCH_POW2;
    s->branch_k = 3;
    s->time_limit_seconds = std::numeric_limits<double>::infinity();
    s->natural_lower_bound = -std::numeric_limits<double>::infinity();
    s->precision = GB_PREC_FP64;
    s->max_nodes = 0;
    s->autotune = 0;
}

extern "C" gb_status gb_branching_dataset(const gb_problem* problem, const gb_objective_desc* objective,
                                          const gb_lmo_desc* lmo, const gb_settings* settings,
                                          int max_nodes, float* features, float* labels, int capacity,
                                          int* count) {
    if (!settings || !features || !labels || !count || capacity < 0 || max_nodes < 0)
        return gb::fail(GB_ERR_INVALID_ARGUMENT, "invalid dataset arguments");
    if (!objective || objective->kind != GB_OBJ_DENSE_QUADRATIC || !lmo || lmo->kind != GB_LMO_BOX)
        return gb::fail(GB_ERR_INVALID_ARGUMENT, "strong-branching labels need a dense objective over a box");
    gb_settings s = *settings;
    s.form = 1;
    // the tree the scorer will be used in: branch_k-variable splits, chosen
    // here by fractionality (top_k_infeasible)
    s.branch_rule = GB_BRANCH_POW2;
    gb::Labeler lab;
    lab.max_nodes = max_nodes;
    lab.features = features;
    lab.labels = labels;
    lab.capacity = capacity;
    gb_result r{};
    gb_status st;
    try {
        st = gb::solve_impl(problem, objective, lmo, &s, &r, &lab);
    } catch (const std::exception& e) {
        return gb::fail(GB_ERR_RUNTIME, e.what());
    }
    *count = lab.count;
    return st;
}

extern "C" gb_status gb_solve(const gb_problem* problem, const gb_objective_desc* objective,
                              const gb_lmo_desc* lmo, const gb_settings* settings,
                              gb_result* result) {
    if (!result) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null result");
    try {
        return gb::solve_impl(problem, objective, lmo, settings, result);
    } catch (const std::bad_alloc&) {
        return gb::fail(GB_ERR_RUNTIME, "host memory exhausted");
    } catch (const std::exception& e) {
        return gb::fail(GB_ERR_RUNTIME, e.what());
    }
}

extern "C" gb_status gb_children_push(gb_children* children, const gb_node* child) {
    if (!children || !child || !child->lb || !child->ub)
        return gb::fail(GB_ERR_INVALID_ARGUMENT, "null child node");
    try {
        gb_children::Child c;
        c.id = child->id;
        c.parent = child->parent_id;
        c.depth = child->depth;
        c.bound = child->dual_bound;
        c.lb.assign(child->lb, child->lb + children->n);
        c.ub.assign(child->ub, child->ub + children->n);
        children->kids.push_back(std::move(c));
        return GB_OK;
    } catch (const std::exception& e) {
        return gb::fail(GB_ERR_RUNTIME, e.what());
    }
}

// The low-level solver object (DeviceResidentHybridSolver,
// device_resident_solver.cuh:1184-1298) over solve_impl: the problem is
// copied, nodes are pushed with the reference's checks, and each solve
// consumes the pushed nodes.
struct gb_solver {
    std::vector<double> lb, ub;
    std::vector<int> iv;
    gb_problem problem{};
    gb_objective_desc objective{};
    gb_lmo_desc lmo{};
    gb_settings settings{};
    std::vector<gb::InitNode> pending;
    gb_scheduling_fn sched = nullptr;
    void* sched_user = nullptr;
    gb_solver_stats stats{};
    std::vector<double> inc_x;
    bool has_inc = false;
};

extern "C" gb_status gb_solver_create(const gb_problem* problem, const gb_objective_desc* objective,
                                      const gb_lmo_desc* lmo, const gb_settings* settings, gb_solver** out) {
    if (!out) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null solver handle");
    *out = nullptr;
    try {
        gb_status st = gb::validate_settings(problem, objective, lmo, settings);
        if (st != GB_OK) return st;
        st = gb::validate_descriptors(objective, &settings->algorithm, lmo, true, settings->plugin);
        if (st != GB_OK) return st;
        auto* sv = new gb_solver();
        const int n = problem->n_vars;
        sv->lb.assign(problem->lb, problem->lb + n);
        sv->ub.assign(problem->ub, problem->ub + n);
        if (problem->n_int_vars > 0) sv->iv.assign(problem->int_vars, problem->int_vars + problem->n_int_vars);
        sv->problem = gb_problem{n, sv->lb.data(), sv->ub.data(), sv->iv.data(), (int)sv->iv.size()};
        sv->objective = *objective;
        sv->lmo = *lmo;
        sv->settings = *settings;
        sv->settings.root_initial_point = nullptr;
        sv->settings.root_initial_point_size = 0;
        sv->stats.incumbent_value = std::numeric_limits<double>::infinity();
        sv->stats.dual_bound = -std::numeric_limits<double>::infinity();
        sv->stats.last_gpu_gap = std::numeric_limits<double>::infinity();
        sv->stats.effective_gpu_batch_size = settings->gpu_batch_size;
        sv->stats.effective_num_streams = settings->num_streams;
        *out = sv;
        return GB_OK;
    } catch (const std::exception& e) {
        return gb::fail(GB_ERR_RUNTIME, e.what());
    }
}

extern "C" gb_status gb_solver_destroy(gb_solver* solver) {
    delete solver;
    return GB_OK;
}

extern "C" gb_status gb_solver_push(gb_solver* solver, const gb_node* node) {
    if (!solver || !node) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null solver or node");
    // device_resident_solver.cuh:1247-1260
    if (!node->lb || !node->ub || node->relaxation_iterations < 0
            || (node->relaxation_iterations > 0 && !node->relaxation_x))
        return gb::fail(GB_ERR_INVALID_ARGUMENT, "invalid node buffer dimensions or resume state");
    const int n = solver->problem.n_vars;
    if (node->relaxation_x)
        for (int j = 0; j < n; ++j)
            if (!std::isfinite(node->relaxation_x[j]) || node->relaxation_x[j] < node->lb[j] - 1e-9
                    || node->relaxation_x[j] > node->ub[j] + 1e-9)
                return gb::fail(GB_ERR_INVALID_ARGUMENT, "initial/resumed node iterate violates finite bounds");
    try {
        gb::InitNode in;
        in.id = node->id;
        in.parent = node->parent_id;
        in.depth = node->depth;
        in.lb.assign(node->lb, node->lb + n);
        in.ub.assign(node->ub, node->ub + n);
        if (node->relaxation_x) in.rx.assign(node->relaxation_x, node->relaxation_x + n);
        in.riters = node->relaxation_iterations;
        in.bound = node->dual_bound;
        solver->pending.push_back(std::move(in));
        return GB_OK;
    } catch (const std::exception& e) {
        return gb::fail(GB_ERR_RUNTIME, e.what());
    }
}

extern "C" gb_status gb_solver_set_scheduling_callback(gb_solver* solver, gb_scheduling_fn fn, void* user) {
    if (!solver) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null solver");
    solver->sched = fn;
    solver->sched_user = user;
    return GB_OK;
}

extern "C" gb_status gb_solver_set_progress_callback(gb_solver* solver, gb_progress_fn fn, void* user) {
    if (!solver) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null solver");
    solver->settings.progress = fn;
    solver->settings.progress_user = user;
    return GB_OK;
}

extern "C" gb_status gb_solver_solve(gb_solver* solver) {
    if (!solver) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null solver");
    try {
        const int n = solver->problem.n_vars;
        std::vector<double> x(n);
        gb_result r{};
        r.x = x.data();
        gb::SolveHooks hk;
        hk.nodes = &solver->pending;
        hk.sched = solver->sched;
        hk.sched_user = solver->sched_user;
        const gb_status st = gb::solve_impl(&solver->problem, &solver->objective, &solver->lmo, &solver->settings,
                                            &r, nullptr, &hk);
        solver->pending.clear();
        if (st != GB_OK) return st;
        gb_solver_stats& S = solver->stats;
        S.incumbent_value = r.objective;
        S.dual_bound = r.dual_bound;
        S.nodes_processed = r.nodes_processed;
        S.last_gpu_iterations = hk.last_iters;
        S.last_gpu_gap = hk.last_gap;
        S.last_gpu_converged = hk.last_conv ? 1 : 0;
        S.reached_tolerance = r.reached_tolerance;
        S.reached_time_limit = r.reached_time_limit;
        S.fw_iterations = r.fw_iterations;
        S.elapsed_seconds = r.elapsed_seconds;
        S.relax_seconds = r.relax_seconds;
        solver->has_inc = std::isfinite(r.objective);
        if (solver->has_inc) solver->inc_x = x;
        else solver->inc_x.clear();
        return GB_OK;
    } catch (const std::bad_alloc&) {
        return gb::fail(GB_ERR_RUNTIME, "host memory exhausted");
    } catch (const std::exception& e) {
        return gb::fail(GB_ERR_RUNTIME, e.what());
    }
}

extern "C" gb_status gb_solver_get_stats(const gb_solver* solver, gb_solver_stats* out) {
    if (!solver || !out) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null solver or output");
    *out = solver->stats;
    return GB_OK;
}

extern "C" gb_status gb_solver_incumbent_x(const gb_solver* solver, double* x, int* has) {
    if (!solver || !has) return gb::fail(GB_ERR_INVALID_ARGUMENT, "null solver or output");
    *has = solver->has_inc ? 1 : 0;
    if (solver->has_inc && x) std::copy(solver->inc_x.begin(), solver->inc_x.end(), x);
    return GB_OK;
}
