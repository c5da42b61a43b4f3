// Host side of the device-resident search (gb_settings.device_frontier):
// buffers, the root, one eager round, then one captured HIP graph per round
// replayed until the search stops.  The host reads the 200-byte search
// state once per group of replays (and on the time limit asks the next
// round to stop); nodes, bounds and iterates never cross the bus.
//
// The rounds reproduce the host driver's one-stream schedule (bnb.cpp with
// num_streams = 1, after device_resident_solver.cuh:1570-2007): pop a batch
// when none is in flight; when a batch completes and a full batch waits,
// pop and relax that one before the completed batch is post-processed; the
// nodes of a batch are post-processed in order.  The same problem and
// settings therefore give the same tree, incumbent, bound and counts as
// gb_solve with num_streams = 1.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "gb_dsearch.h"
#include "gb_host.h"

namespace gb {

namespace {

constexpr double kInfH = std::numeric_limits<double>::infinity();

// One device allocation carved into aligned arrays.
class Arena {
  public:
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    template <class T>
    void plan(T** slot, size_t count) {
        items_.push_back({reinterpret_cast<void**>(slot), count * sizeof(T)});
    }
    gb_status commit() {
        size_t total = 0;
        for (auto& it : items_) total += (it.bytes + 255) & ~size_t(255);
        GB_HIP(hipMalloc(&base_, total ? total : 256));
        char* p = static_cast<char*>(base_);
        for (auto& it : items_) {
            *it.slot = p;
            p += (it.bytes + 255) & ~size_t(255);
        }
        return GB_OK;
    }

  private:
    struct Item {
        void** slot;
        size_t bytes;
    };
    std::vector<Item> items_;
    void* base_ = nullptr;
};

void plan_batch(Arena& a, DsBatch& T, int B, int n) {
    const size_t bn = (size_t)B * n;
    a.plan(&T.count, 16);
    a.plan(&T.relax_count, 16);
    a.plan(&T.lb, bn);
    a.plan(&T.ub, bn);
    a.plan(&T.x, bn);
    a.plan(&T.bv, bn);
    a.plan(&T.eps, B);
    a.plan(&T.gap, B);
    a.plan(&T.val, B);
    a.plan(&T.bvv, B);
    a.plan(&T.cut, B);
    a.plan(&T.bound, B);
    a.plan(&T.it, B);
    a.plan(&T.st, B);
    a.plan(&T.id, B);
    a.plan(&T.parent, B);
    a.plan(&T.depth, B);
    a.plan(&T.riters, B);
    a.plan(&T.b8, 2 * bn);
}

struct StreamHolder {
    hipStream_t s = nullptr;
    ~StreamHolder() {
        if (s) (void)hipStreamDestroy(s);
    }
};
struct GraphHolder {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~GraphHolder() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
};
struct PinnedState {
    DsState* p = nullptr;
    int* flag = nullptr;
    ~PinnedState() {
        if (p) (void)hipHostFree(p);
        if (flag) (void)hipHostFree(flag);
    }
};

}  // namespace

gb_status device_search(const gb_problem* pr, const gb_objective_desc* o, const gb_lmo_desc* l,
                        const gb_settings* s, gb_result* res) {
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    const int n = pr->n_vars;
    const int B = s->gpu_batch_size;
    if (s->form != 1 && s->form != 2)
        return fail(GB_ERR_INVALID_ARGUMENT, "device_frontier runs the batched forms (form 1 or 2)");
    if (o->kind != GB_OBJ_DENSE_QUADRATIC || (l->kind != GB_LMO_BOX && !fast_simplex_ok(l)) || n > kMaxFastN)
        return fail(GB_ERR_INVALID_ARGUMENT,
                    "device_frontier needs a dense quadratic over a box or a simplex with an integral right side, "
                    "n <= 512");
    if (pr->n_int_vars > n) return fail(GB_ERR_INVALID_ARGUMENT, "device_frontier needs distinct integer variables");
    if (B < 1 || B > kDsMaxBatch)
        return fail(GB_ERR_INVALID_ARGUMENT, "device_frontier needs 1 <= gpu_batch_size <= 1024");
    if (s->branch_rule != GB_BRANCH_POW2 && s->branch_rule != GB_BRANCH_MOST_FRACTIONAL)
        return fail(GB_ERR_INVALID_ARGUMENT, "device_frontier branches by pow2 or most-fractional");
    if (s->branch_rule == GB_BRANCH_POW2 && (s->branch_k < 1 || s->branch_k > kDsMaxK))
        return fail(GB_ERR_INVALID_ARGUMENT, "device_frontier supports branch_k 1..3");
    for (int j = 0; j < n; ++j) {
        const bool l01 = pr->lb[j] == 1.0 || (pr->lb[j] == 0.0 && !std::signbit(pr->lb[j]));
        const bool u01 = pr->ub[j] == 1.0 || (pr->ub[j] == 0.0 && !std::signbit(pr->ub[j]));
        if (!l01 || !u01) return fail(GB_ERR_INVALID_ARGUMENT, "device_frontier needs 0/1 variable bounds");
    }
    int device = 0;
    GB_HIP(hipGetDevice(&device));
    ensure_shutdown_hook();

    // Node slots: the frontier of the largest searches holds millions of
    // nodes (about 10 n + 53 bytes each); a quarter of free memory, at most
    // 32 GiB and 16 M slots, or what a node cap can need.
    size_t free_b = 0, total_b = 0;
    GB_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per_slot = (size_t)10 * n + 53;
    size_t cap = std::min(std::min(free_b / 4, (size_t)32 << 30) / per_slot, (size_t)16 << 20);
    if (s->max_nodes > 0) cap = std::min(cap, (size_t)(9 * (s->max_nodes + B) + 16));
    if (s->device_nodes > 0) cap = (size_t)s->device_nodes;
    cap = std::min(cap, (size_t)1 << 30);
    cap = std::max(cap, (size_t)16 * B + 16);
    const int C = (int)cap;

    DsLaunch L{};
    L.n = n;
    L.batch = B;
    L.n_iv = pr->n_int_vars;
    L.int_tol = s->integer_tol;
    L.abs_tol = s->abs_gap_tol;
    L.rel_tol = s->rel_gap_tol;
    L.rule = s->branch_rule;
    L.branch_k = s->branch_rule == GB_BRANCH_POW2 ? s->branch_k : 1;
    L.child_warm = l->kind == GB_LMO_SIMPLEX ? 0 : 1;
    L.store.capacity = C;
    Arena arena;
    arena.plan(&L.S, 1);
    int* iv_d = nullptr;
    double* eps_d = nullptr;
    uint8_t* flags_d = nullptr;
    int* counters = nullptr;
    arena.plan(&iv_d, std::max(1, pr->n_int_vars));
    arena.plan(&eps_d, n + 2);
    arena.plan(&flags_d, n);
    arena.plan(&counters, 64);
    plan_batch(arena, L.A, B, n);
    plan_batch(arena, L.B, B, n);
    arena.plan(&L.post.nb, B);
    arena.plan(&L.post.cand, B);
    arena.plan(&L.post.kind, B);
    arena.plan(&L.post.nvars, B);
    arena.plan(&L.post.vars, (size_t)B * kDsMaxK);
    arena.plan(&L.post.pushes, B);
    arena.plan(&L.post.push_off, B);
    arena.plan(&L.post.id_base, B);
    arena.plan(&L.post.slot_base, B);
    arena.plan(&L.post.fam_row, B);
    arena.plan(&L.post.fam_off, B);
    arena.plan(&L.post.fresh, (size_t)8 * B);
    arena.plan(&L.post.sorted, (size_t)8 * B);
    arena.plan(&L.post.inc_x, n);
    arena.plan(&L.store.b8, (size_t)C * 2 * n);
    arena.plan(&L.store.x0, (size_t)C * n);
    arena.plan(&L.store.id, C);
    arena.plan(&L.store.parent, C);
    arena.plan(&L.store.depth, C);
    arena.plan(&L.store.riters, C);
    arena.plan(&L.store.warm, C);
    arena.plan(&L.store.free_list, C);
    arena.plan(&L.store.f[0], C);
    arena.plan(&L.store.f[1], C);
    gb_status st = arena.commit();
    if (st != GB_OK) return st;
    L.iv = iv_d;
    L.eps_table = eps_d;

    StreamHolder sh;
    GB_HIP(hipStreamCreateWithFlags(&sh.s, hipStreamNonBlocking));
    hipStream_t stream = sh.s;

    // Root, tolerances, flags, the free list and the state.
    {
        std::vector<double> eps(n + 2);
        for (int d = 0; d < n + 2; ++d) eps[d] = depth_tolerance(s->gpu_eps, s->gpu_eps_decay, s->gpu_eps_min, d);
        GB_HIP(hipMemcpy(eps_d, eps.data(), eps.size() * 8, hipMemcpyHostToDevice));
        if (pr->n_int_vars > 0)
            GB_HIP(hipMemcpy(iv_d, pr->int_vars, (size_t)pr->n_int_vars * 4, hipMemcpyHostToDevice));
        std::vector<uint8_t> fl(n, 0);
        for (int k = 0; k < pr->n_int_vars; ++k) fl[pr->int_vars[k]] = 1;
        GB_HIP(hipMemcpy(flags_d, fl.data(), n, hipMemcpyHostToDevice));
        GB_HIP(hipMemset(counters, 0, 64 * sizeof(int)));
        std::vector<uint8_t> b8(2 * n);
        for (int j = 0; j < n; ++j) {
            b8[j] = pr->lb[j] == 1.0 ? 1 : 0;
            b8[n + j] = pr->ub[j] == 1.0 ? 1 : 0;
        }
        GB_HIP(hipMemcpy(L.store.b8, b8.data(), 2 * n, hipMemcpyHostToDevice));
        const bool has_rx = s->root_initial_point && s->root_initial_point_size == n;
        if (has_rx) GB_HIP(hipMemcpy(L.store.x0, s->root_initial_point, (size_t)n * 8, hipMemcpyHostToDevice));
        const int zero = 0, minus = -1;
        const uint8_t warm = has_rx ? 1 : 0;
        GB_HIP(hipMemcpy(L.store.id, &zero, 4, hipMemcpyHostToDevice));
        GB_HIP(hipMemcpy(L.store.parent, &minus, 4, hipMemcpyHostToDevice));
        GB_HIP(hipMemcpy(L.store.depth, &zero, 4, hipMemcpyHostToDevice));
        GB_HIP(hipMemcpy(L.store.riters, &zero, 4, hipMemcpyHostToDevice));
        GB_HIP(hipMemcpy(L.store.warm, &warm, 1, hipMemcpyHostToDevice));
        const DsEntry root{s->natural_lower_bound, 0, 0};
        GB_HIP(hipMemcpy(L.store.f[0], &root, sizeof(root), hipMemcpyHostToDevice));
        const int e = ds_launch_init_free(L.store.free_list, C, stream);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        DsState S0{};
        S0.inc = kInfH;
        S0.final_bound = -kInfH;
        S0.max_nodes = s->max_nodes;
        S0.next_id = 1;
        S0.fsize = 1;
        S0.free_top = C - 1;  // slot 0 holds the root
        S0.inc_row = -1;
        GB_HIP(hipMemcpyAsync(L.S, &S0, sizeof(S0), hipMemcpyHostToDevice, stream));
        GB_HIP(hipStreamSynchronize(stream));
    }

    std::unique_ptr<gb_mixed_q, gb_status (*)(gb_mixed_q*)> mq(nullptr, gb_mixed_destroy);
    if (s->form == 2) {
        gb_mixed_q* h = nullptr;
        st = gb_mixed_prepare(o, s->precision == GB_PREC_BF16 ? 1 : 3, stream, &h);
        if (st != GB_OK) return st;
        mq.reset(h);
    }
    const gb_mixed_options mo{0, 0, 0, 1e-6f, 0, nullptr, nullptr, nullptr};
    const gb_fast_options fo{kFastRefresh, 0, 0, nullptr, nullptr};
    auto relax = [&](const DsBatch& T, int* counter) -> gb_status {
        if (s->form == 1)
            return relax_batch_fast(o, &s->algorithm, l, B, s->gpu_max_iters, T.x, T.lb, T.ub, T.eps, T.it, T.st,
                                    T.gap, T.val, T.bv, T.bvv, flags_d, s->integer_tol, T.cut, &fo, stream, counter,
                                    false, T.relax_count);
        return relax_mixed_impl(mq.get(), o, &s->algorithm, l, B, s->gpu_max_iters, T.x, T.lb, T.ub, T.eps, T.it,
                                T.st, T.gap, T.val, T.bv, T.bvv, flags_d, s->integer_tol, T.cut, &mo, stream,
                                counter, T.relax_count);
    };
    auto round = [&]() -> gb_status {
        int e = ds_launch_round_begin(L, stream);
        if (e == 0) e = ds_launch_select(L, 0, stream);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        gb_status r = relax(L.A, counters);
        if (r != GB_OK) return r;
        e = ds_launch_select(L, 1, stream);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        r = relax(L.B, counters + 16);
        if (r != GB_OK) return r;
        e = ds_launch_post(L, stream);
        if (e == 0) e = ds_launch_round_end(L, stream);
        if (e != 0) return hip_fail(static_cast<hipError_t>(e), __FILE__, __LINE__);
        return GB_OK;
    };

    PinnedState hs;
    GB_HIP(hipHostMalloc(reinterpret_cast<void**>(&hs.p), sizeof(DsState), hipHostMallocDefault));
    GB_HIP(hipHostMalloc(reinterpret_cast<void**>(&hs.flag), sizeof(int), hipHostMallocDefault));
    *hs.flag = 1;
    auto read_state = [&]() -> gb_status {
        GB_HIP(hipMemcpyAsync(hs.p, L.S, sizeof(DsState), hipMemcpyDeviceToHost, stream));
        GB_HIP(hipStreamSynchronize(stream));
        return GB_OK;
    };
    // The first rounds eagerly (they also settle every launch's one-time
    // set-up, and small searches end within them), then the round as one
    // graph.
    constexpr int kEagerRounds = 2;
    for (int r = 0; r < kEagerRounds; ++r) {
        st = round();
        if (st != GB_OK) return st;
        st = read_state();
        if (st != GB_OK) return st;
        if (hs.p->finished) break;
    }
    GraphHolder gh;
    if (!hs.p->finished) {
        GB_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        st = round();
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(stream, &g);
        gh.g = g;
        if (st != GB_OK) return st;
        GB_HIP(ce);
        GB_HIP(hipGraphInstantiate(&gh.e, gh.g, nullptr, nullptr, 0));
    }
    int group = 1;
    bool stop_sent = false;
    while (!hs.p->finished) {
        for (int r = 0; r < group; ++r) GB_HIP(hipGraphLaunch(gh.e, stream));
        st = read_state();
        if (st != GB_OK) return st;
        if (group < 16) group *= 2;
        const double el = std::chrono::duration<double>(Clock::now() - t0).count();
        if (!hs.p->finished && el >= s->time_limit_seconds && !stop_sent) {
            // the next round stops at its start, as the host loop does
            GB_HIP(hipMemcpyAsync(reinterpret_cast<char*>(L.S) + offsetof(DsState, stop_request), hs.flag,
                                  sizeof(int), hipMemcpyHostToDevice, stream));
            stop_sent = true;
            group = 1;
        }
    }
    const DsState& S = *hs.p;
    res->objective = S.inc;
    res->dual_bound = S.final_bound;
    res->nodes_processed = (int)S.done_nodes;
    res->reached_tolerance = S.tol_hit;
    res->reached_time_limit = S.limit_hit;
    res->fw_iterations = S.fw_iters;
    if (res->x) {
        if (S.inc < kInfH) {
            GB_HIP(hipMemcpy(res->x, L.post.inc_x, (size_t)n * 8, hipMemcpyDeviceToHost));
        } else {
            for (int j = 0; j < n; ++j) res->x[j] = std::nan("");
        }
    }
    res->elapsed_seconds = std::chrono::duration<double>(Clock::now() - t0).count();
    res->relax_seconds = res->elapsed_seconds;  // every relaxation and the tree work run on the device
    if (S.overflow)
        return fail(GB_ERR_RUNTIME, "device_frontier: node slots exhausted (the result holds the bound so far)");
    return GB_OK;
}

}  // namespace gb
