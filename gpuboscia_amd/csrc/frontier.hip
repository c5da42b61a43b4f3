// Device-resident branch-and-bound rounds for gfx950: selection of the next
// node batch from the frontier, post-processing of a relaxed batch
// (certificate, incumbent, pruning, stop test, branching, children), and
// the merge of the new entries into the frontier.  The host driver's
// sequential pass over a batch (bnb.cpp, after device_resident_solver.cuh
// :1900-2007) is evaluated here with scans: the incumbent after each node is
// a prefix minimum, the live bound a combination of prefix and suffix
// minima, so every decision equals the one the host makes at that node.
//
// The frontier is one sorted array of (bound, id, slot) entries (two
// buffers): a selection takes a prefix, the entries of a round are sorted by
// family (a node's children share its bound and carry consecutive ids) and
// merged in by rank, every element placed by two binary searches.  Stale
// entries (bound at or above the incumbent) stay until a selection drains
// them, as in the host's heap, so the frontier size that decides a refill is
// the host's.
#include "../../include/gpuboscia_amd.h"
#include "gb_device.h"
#include "gb_dsearch.h"
#include "gb_stdsort.h"

namespace gb {
namespace {

constexpr double kInfD = __builtin_inf();

// gap_utils.h:14-62 (the host restatement in bnb.cpp), with std::max /
// std::min written out so that they pick the same operand.
__device__ inline double ds_certified(double inherited, double primal, double gap) {
    if (!isfinite(primal) || !isfinite(gap) || gap < 0.0) return inherited;
    const double v = primal - gap;
    return inherited < v ? v : inherited;
}
__device__ inline double ds_min(double a, double b) { return b < a ? b : a; }
__device__ inline bool ds_within_tolerance(double bound, double inc, double at, double rt) {
    if (!isfinite(bound) || !isfinite(inc)) return false;
    const double ga = (inc >= kInfD || bound <= -kInfD) ? kInfD : fabs(inc - bound);
    double gr = kInfD;
    const double fb = fabs(bound), fi = fabs(inc);
    const double scale = fi < fb ? fi : fb;
    if (scale != 0.0) gr = (inc - bound) / scale;
    return ga <= at || gr <= rt;
}

__device__ inline bool ds_less(const DsEntry& a, const DsEntry& b) {
    return a.bound < b.bound || (a.bound == b.bound && a.id < b.id);
}

__global__ void ds_round_begin_kernel(DsLaunch L) {
    DsState* S = L.S;
    S->posted = 0;
    if (S->finished) return;
    S->refilled = 0;
    S->inc_row = -1;
    S->pushes = 0;
    S->families = 0;
    S->rounds += 1;
    for (int m = 0; m < 2; ++m) {
        S->plan_take[m] = 0;
        S->plan_drop[m] = 0;
    }
    *L.B.relax_count = 0;
    if (S->stop_request || (S->max_nodes > 0 && S->done_nodes >= S->max_nodes)) {
        // the host's loop-top stop: the live bound over the frontier and the
        // batch in flight
        double b = S->inc;
        if (S->fsize > 0) b = ds_min(b, L.store.f[S->fbuf][S->fhead].bound);
        if (S->a_live) {
            const int N = *L.A.count;
            double m = kInfD;
            for (int i = N - 1; i >= 0; --i) m = ds_min(m, L.A.bound[i]);
            if (N > 0) b = ds_min(b, m);
        }
        S->final_bound = b;
        S->limit_hit = 1;
        S->finished = 1;
        *L.A.relax_count = 0;
    }
}

// Pops the next batch (m = 0: into A when nothing is in flight; m = 1: the
// refill of B, only when at least a full batch waits).  Entries whose bound
// reaches the incumbent are dropped on the way, so a batch that cannot be
// filled drains the frontier.
__global__ void ds_select_kernel(DsLaunch L, int m) {
    DsState* S = L.S;
    DsBatch T = m ? L.B : L.A;
    if (S->finished || (m == 0 && S->a_live) || (m == 1 && S->fsize < L.batch)) {
        *T.relax_count = 0;
        return;
    }
    const DsEntry* q = L.store.f[S->fbuf] + S->fhead;
    const double inc = S->inc;
    int lo = 0, hi = S->fsize;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(q[mid].bound >= inc)) lo = mid + 1;
        else hi = mid;
    }
    const int take = lo < L.batch ? lo : L.batch;
    int drop = 0;
    S->plan_first[m] = S->fhead;
    if (lo >= L.batch) {
        S->fhead += L.batch;
        S->fsize -= L.batch;
    } else {
        drop = S->fsize - take;
        S->fhead += S->fsize;
        S->fsize = 0;
    }
    S->plan_take[m] = take;
    S->plan_drop[m] = drop;
    S->plan_free_base[m] = S->free_top;
    S->free_top += take + drop;
    *T.count = take;
    *T.relax_count = take;
    if (m == 0) {
        S->a_live = take > 0 ? 1 : 0;
        if (take == 0) {
            S->finished = 1;
            S->final_bound = inc;
        }
    } else {
        S->refilled = take > 0 ? 1 : 0;
    }
}

// Row r of the selected batch: bounds, start point, tolerance, cutoff; the
// node's slot goes back to the free list.
__global__ void ds_stage_kernel(DsLaunch L, int m) {
    DsState* S = L.S;
    const int r = blockIdx.x;
    if (r >= S->plan_take[m]) return;
    DsBatch T = m ? L.B : L.A;
    const int n = L.n;
    const DsEntry e = L.store.f[S->fbuf][S->plan_first[m] + r];
    const int s = e.slot;
    const uint8_t* b8 = L.store.b8 + (size_t)s * 2 * n;
    uint8_t* tb8 = T.b8 + (size_t)r * 2 * n;
    const bool warm = L.store.warm[s] != 0;
    const double* x0 = L.store.x0 + (size_t)s * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const uint8_t l = b8[j], u = b8[n + j];
        tb8[j] = l;
        tb8[n + j] = u;
        T.lb[(size_t)r * n + j] = l;
        T.ub[(size_t)r * n + j] = u;
        T.x[(size_t)r * n + j] = warm ? x0[j] : (double)l;
    }
    if (threadIdx.x == 0) {
        const int depth = L.store.depth[s];
        T.id[r] = L.store.id[s];
        T.parent[r] = L.store.parent[s];
        T.depth[r] = depth;
        T.riters[r] = L.store.riters[s];
        T.bound[r] = e.bound;
        T.it[r] = L.store.riters[s];
        T.st[r] = warm ? GB_NODE_INITIAL_POINT_SUPPLIED : 0;
        T.eps[r] = L.eps_table[depth < n + 1 ? depth : n + 1];
        T.cut[r] = S->inc;
        T.bvv[r] = 1e300;
        L.store.free_list[S->plan_free_base[m] + r] = s;
    }
}

__global__ void ds_free_dropped_kernel(DsLaunch L, int m) {
    DsState* S = L.S;
    const int drop = S->plan_drop[m];
    if (drop == 0) return;
    const int take = S->plan_take[m];
    const DsEntry* q = L.store.f[S->fbuf] + S->plan_first[m] + take;
    int* fl = L.store.free_list + S->plan_free_base[m] + take;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < drop; i += gridDim.x * blockDim.x) fl[i] = q[i].slot;
}

// One wave per row of A: certified bound, incumbent candidate, kind, and
// the branching variables of a fractional node (top_k_infeasible or the
// most fractional variable, branching.h:24-65, as the host computes them).
__global__ void __launch_bounds__(64) ds_post_node_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (S->finished || !S->a_live) return;
    const int r = blockIdx.x;
    const DsBatch A = L.A;
    if (r >= *A.count) return;
    struct Cand {
        int idx;
        double dist;
    };
    __shared__ Cand cand[512];
    const int lane = threadIdx.x;
    const double gap = A.gap[r];
    const uint8_t st = A.st[r];
    const bool cert = isfinite(gap) && gap >= 0.0 && gap < A.eps[r];
    const int kind = st == GB_NODE_INFEASIBLE ? kDsInfeasible
                     : st >= GB_NODE_INTEGRAL_UNCONVERGED ? (cert ? kDsIntegral : kDsRequeue)
                                                          : kDsBranch;
    if (lane == 0) {
        const double bvv = A.bvv[r];
        L.post.nb[r] = ds_certified(A.bound[r], A.val[r], gap);
        L.post.cand[r] = (kind != kDsInfeasible && bvv < 1e299 && isfinite(bvv)) ? bvv : kInfD;
        L.post.kind[r] = kind;
        L.post.nvars[r] = 0;
    }
    if (kind != kDsBranch) return;
    // fractional candidates in the order of the integer variables
    const int n = L.n;
    const double* x = A.x + (size_t)r * n;
    int count = 0;
    for (int c0 = 0; c0 < L.n_iv; c0 += kWave) {
        const int p = c0 + lane;
        bool ok = false;
        double d = 0.0;
        int v = 0;
        if (p < L.n_iv) {
            v = L.iv[p];
            const double xv = x[v];
            const double frac = xv - floor(xv);
            const double up = 1.0 - frac;
            d = up < frac ? up : frac;
            ok = d > L.int_tol;
        }
        const unsigned long long mask = __ballot(ok);
        if (ok) cand[count + __popcll(mask & ((1ull << lane) - 1ull))] = Cand{v, d};
        count += __popcll(mask);
    }
    __syncthreads();
    if (lane != 0) return;
    int* out = L.post.vars + (size_t)r * kDsMaxK;
    int k = 0;
    if (L.rule == GB_BRANCH_MOST_FRACTIONAL) {
        double bd = -1.0;
        int best = -1;
        for (int c = 0; c < count; ++c)
            if (cand[c].dist > bd) {
                bd = cand[c].dist;
                best = cand[c].idx;
            }
        if (best >= 0) out[k++] = best;
    } else {
        // the k largest distances, in order; when equal distances decide
        // the prefix, std::sort's order of the whole candidate list
        const int K = L.branch_k;
        int best_i[kDsMaxK];
        double best_d[kDsMaxK];
        int m = 0;
        for (int c = 0; c < count; ++c) {
            const double dist = cand[c].dist;
            if (m < K || dist > best_d[m - 1]) {
                int p = m < K ? m++ : m - 1;
                while (p > 0 && best_d[p - 1] < dist) {
                    best_d[p] = best_d[p - 1];
                    best_i[p] = best_i[p - 1];
                    --p;
                }
                best_d[p] = dist;
                best_i[p] = cand[c].idx;
            }
        }
        bool tie = false;
        for (int c = 1; c < m && !tie; ++c) tie = best_d[c] == best_d[c - 1];
        if (!tie && m == K)
            for (int c = 0; c < count; ++c)
                if (cand[c].dist == best_d[m - 1] && cand[c].idx != best_i[m - 1]) {
                    tie = true;
                    break;
                }
        if (tie) {
            std_order_sort(cand, cand + count, [](const Cand& a, const Cand& b) { return a.dist > b.dist; });
            const int mm = count < K ? count : K;
            for (int c = 0; c < mm; ++c) out[k++] = cand[c].idx;
        } else {
            for (int c = 0; c < m; ++c) out[k++] = best_i[c];
        }
    }
    L.post.nvars[r] = k;
}

// Inclusive block scans over up to 1024 threads (one element each).
template <class T, class Op>
__device__ T block_scan(T v, T* buf, Op op) {
    const int t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const T o = t >= off ? buf[t - off] : v;
        __syncthreads();
        if (t >= off) v = op(o, v);
        buf[t] = v;
        __syncthreads();
    }
    return v;
}

// The host's sequential pass over the rows of A (bnb.cpp), by scans: the
// incumbent after each row, prune / requeue / branch, the stop test after
// each finished row, and the ids, slots and entry offsets of every push.
__global__ void __launch_bounds__(1024) ds_post_scan_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (S->finished || !S->a_live) return;
    const DsBatch A = L.A;
    const int N = *A.count;
    const int t = threadIdx.x;
    const bool row = t < N;
    __shared__ double dbuf[1024];
    __shared__ long long lbuf[1024];
    __shared__ int first_stop;
    __shared__ int last_improve;
    if (t == 0) {
        first_stop = 1 << 30;
        last_improve = -1;
    }
    const double inc0 = S->inc;
    const double cand = row ? L.post.cand[t] : kInfD;
    const double nb = row ? L.post.nb[t] : 0.0;
    const int kind = row ? L.post.kind[t] : kDsInfeasible;
    const int nv = row ? L.post.nvars[t] : 0;
    const double launch_bound = row ? A.bound[t] : kInfD;
    auto dmin = [](double a, double b) { return ds_min(a, b); };
    auto lsum = [](long long a, long long b) { return a + b; };
    // incumbent after row t (each row's best vertex offered in row order;
    // infeasible rows offer none)
    const double incp = ds_min(inc0, block_scan(cand, dbuf, dmin));
    __syncthreads();
    const double inc_before = t == 0 ? inc0 : dbuf[t - 1] < inc0 ? dbuf[t - 1] : inc0;
    __syncthreads();
    int processed = 0, push = 0, branched = 0;
    if (row) {
        if (kind == kDsInfeasible) {
            processed = 1;
        } else if (nb >= incp) {
            processed = 1;  // pruned
        } else if (kind == kDsRequeue) {
            push = 1;       // continuation: back to the frontier
        } else if (kind == kDsBranch) {
            processed = 1;
            branched = 1;
            push = nv > 0 ? (1 << nv) : 0;
        } else {
            processed = 1;  // integral and converged
        }
    }
    const long long push_incl = block_scan((long long)push, lbuf, lsum);
    __syncthreads();
    const double ftop = S->fsize > 0 ? L.store.f[S->fbuf][S->fhead].bound : kInfD;
    const double qmin = ds_min(ftop, block_scan(push > 0 ? nb : kInfD, dbuf, dmin));
    __syncthreads();
    // minimum launch bound of the rows after t (suffix minimum)
    dbuf[t] = launch_bound;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const double o = t + off < 1024 ? dbuf[t + off] : kInfD;
        __syncthreads();
        dbuf[t] = ds_min(o, dbuf[t]);
        __syncthreads();
    }
    const double after = t + 1 < 1024 ? dbuf[t + 1] : kInfD;
    __syncthreads();
    double refill_min = kInfD;
    int elsewhere = 0;
    if (S->refilled) {
        elsewhere = *L.B.count;
        for (int i = elsewhere - 1; i >= 0; --i) refill_min = ds_min(refill_min, L.B.bound[i]);
    }
    // the stop test after each finished row (finish_check)
    int stop = 0;
    double stop_bound = 0.0;
    if (row && processed) {
        const int remaining = N - t - 1 + elsewhere;
        const bool q_empty = S->fsize == 0 && push_incl == 0;
        if (q_empty && remaining == 0) {
            stop = 1;
            stop_bound = incp;
        } else {
            double b = incp;
            if (!q_empty) b = ds_min(b, qmin);
            if (t + 1 < N) b = ds_min(b, after);
            if (elsewhere > 0) b = ds_min(b, refill_min);
            if (ds_within_tolerance(b, incp, L.abs_tol, L.rel_tol)) {
                stop = 2;
                stop_bound = b;
            }
        }
        if (stop) atomicMin(&first_stop, t);
    }
    __syncthreads();
    const int last = first_stop < N ? first_stop : N - 1;
    const bool eff = row && t <= last;
    if (eff && cand < inc_before) atomicMax(&last_improve, t);
    const int peff = eff ? push : 0;
    const long long off_incl = block_scan((long long)peff, lbuf, lsum);
    __syncthreads();
    const long long br_incl = block_scan((long long)(eff ? branched : 0), lbuf, lsum);
    __syncthreads();
    const long long done_incl = block_scan((long long)(eff ? processed : 0), lbuf, lsum);
    __syncthreads();
    const long long fam_incl = block_scan((long long)(peff > 0 ? 1 : 0), lbuf, lsum);
    __syncthreads();
    const long long it_incl = block_scan((long long)(row ? A.it[t] - A.riters[t] : 0), lbuf, lsum);
    __syncthreads();
    // totals live in thread N - 1's inclusive scans (every scan stops growing past the last row)
    __shared__ long long tot[6];
    __shared__ double q_last;
    if (t == last) q_last = qmin;
    if (t == (N > 0 ? N - 1 : 0)) {
        tot[0] = off_incl;
        tot[1] = br_incl;
        tot[2] = done_incl;
        tot[3] = fam_incl;
        tot[4] = it_incl;
    }
    __syncthreads();
    const int P = (int)tot[0];
    const int free_top = S->free_top;
    const bool fits = P <= free_top && (long long)S->fsize + P <= (long long)L.store.capacity;
    if (row) {
        L.post.pushes[t] = fits ? peff : 0;
        L.post.push_off[t] = (int)(off_incl - peff);
        L.post.id_base[t] = S->next_id + 8 * (int)(br_incl - (eff ? branched : 0));
        L.post.slot_base[t] = free_top - P + (int)(off_incl - peff);
        if (peff > 0) L.post.fam_row[fam_incl - 1] = t;
    }
    if (stop && t == first_stop) {
        S->final_bound = stop_bound;
        S->tol_hit = stop == 2 ? 1 : 0;
    }
    __syncthreads();
    if (t == 0) {
        S->fw_iters += tot[4];
        S->done_nodes += tot[2];
        S->next_id += 8 * (int)tot[1];
        if (last_improve >= 0) {
            S->inc_row = last_improve;
            S->inc = L.post.cand[last_improve];
        }
        S->posted = 1;
        if (first_stop < N) {
            S->finished = 1;
        } else if (!fits) {
            // out of node slots or frontier entries: stop with the bound of
            // every open node (frontier, the entries not pushed, the refill)
            double b = ds_min(S->inc, q_last);
            b = ds_min(b, refill_min);
            S->final_bound = b;
            S->overflow = 1;
            S->finished = 1;
        } else {
            S->free_top = free_top - P;
            S->pushes = P;
            S->families = (int)tot[3];
        }
    }
}

// Children and continuations into their slots and entries; the incumbent's
// best vertex.
__global__ void ds_post_emit_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (!S->posted) return;
    const DsBatch A = L.A;
    const int r = blockIdx.x;
    if (r >= *A.count) return;
    const int n = L.n;
    if (r == S->inc_row)
        for (int j = threadIdx.x; j < n; j += blockDim.x) L.post.inc_x[j] = A.bv[(size_t)r * n + j];
    if (S->finished) return;
    const int P = L.post.pushes[r];
    if (P == 0) return;
    const int kind = L.post.kind[r];
    const int base = L.post.slot_base[r];
    const int off = L.post.push_off[r];
    const double nb = L.post.nb[r];
    const uint8_t* pb8 = A.b8 + (size_t)r * 2 * n;
    const double* px = A.x + (size_t)r * n;
    if (kind == kDsRequeue) {
        const int s = L.store.free_list[base];
        uint8_t* b8 = L.store.b8 + (size_t)s * 2 * n;
        double* x0 = L.store.x0 + (size_t)s * n;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            b8[j] = pb8[j];
            b8[n + j] = pb8[n + j];
            x0[j] = px[j];
        }
        if (threadIdx.x == 0) {
            L.store.id[s] = A.id[r];
            L.store.parent[s] = A.parent[r];
            L.store.depth[s] = A.depth[r];
            L.store.riters[s] = A.it[r];
            L.store.warm[s] = 1;
            L.post.fresh[off] = DsEntry{nb, A.id[r], s};
        }
        return;
    }
    const int k = L.post.nvars[r];
    const int* vars = L.post.vars + (size_t)r * kDsMaxK;
    const int id0 = L.post.id_base[r];
    for (int c = 0; c < P; ++c) {
        const int s = L.store.free_list[base + c];
        uint8_t* b8 = L.store.b8 + (size_t)s * 2 * n;
        double* x0 = L.store.x0 + (size_t)s * n;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            uint8_t l = pb8[j], u = pb8[n + j];
            for (int b = 0; b < k; ++b)
                if (vars[b] == j) {
                    if ((c >> b) & 1) l = 1;  // branched values are fractional: ceil 1, floor 0
                    else u = 0;
                }
            b8[j] = l;
            b8[n + j] = u;
            // the parent's iterate clipped to the child's box
            const double lo = l, hi = u;
            double v = px[j] < lo ? lo : px[j];
            v = hi < v ? hi : v;
            x0[j] = v;
        }
        if (threadIdx.x == 0) {
            L.store.id[s] = id0 + c;
            L.store.parent[s] = A.id[r];
            L.store.depth[s] = A.depth[r] + 1;
            L.store.riters[s] = 0;
            L.store.warm[s] = L.child_warm ? 1 : 0;
            L.post.fresh[off + c] = DsEntry{nb, id0 + c, s};
        }
    }
}

// The families (the entries of one row: a node's children, or one
// continuation) sorted by (bound, first id) with a bitonic sort in LDS, and
// the output offset of each.
__global__ void __launch_bounds__(1024) ds_sort_families_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (!S->posted || S->finished) return;
    const int F = S->families;
    if (F == 0) return;
    __shared__ double kb[1024];
    __shared__ int kid[1024], krow[1024];
    const int t = threadIdx.x;
    int P2 = 1;
    while (P2 < F) P2 <<= 1;
    if (t < P2) {
        if (t < F) {
            const int r = L.post.fam_row[t];
            kb[t] = L.post.nb[r];
            kid[t] = L.post.fresh[L.post.push_off[r]].id;
            krow[t] = r;
        } else {
            kb[t] = kInfD;
            kid[t] = 0x7fffffff;
            krow[t] = -1;
        }
    }
    __syncthreads();
    for (int size = 2; size <= P2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (t < P2) {
                const int o = t ^ stride;
                if (o > t) {
                    const bool asc = (t & size) == 0;
                    const bool gt = kb[t] > kb[o] || (kb[t] == kb[o] && kid[t] > kid[o]);
                    if (gt == asc) {
                        const double b = kb[t];
                        kb[t] = kb[o];
                        kb[o] = b;
                        const int i = kid[t];
                        kid[t] = kid[o];
                        kid[o] = i;
                        const int rr = krow[t];
                        krow[t] = krow[o];
                        krow[o] = rr;
                    }
                }
            }
            __syncthreads();
        }
    // exclusive offsets of the sorted families
    __shared__ int cnt[1024];
    cnt[t] = t < F ? L.post.pushes[krow[t]] : 0;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int o = t >= off ? cnt[t - off] : 0;
        __syncthreads();
        cnt[t] += o;
        __syncthreads();
    }
    if (t < F) {
        L.post.fam_row[t] = krow[t];
        L.post.fam_off[t] = cnt[t] - L.post.pushes[krow[t]];
    }
}

__global__ void ds_expand_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (!S->posted || S->finished) return;
    const int f = blockIdx.x;
    if (f >= S->families) return;
    const int r = L.post.fam_row[f];
    const int cnt = L.post.pushes[r];
    const DsEntry* src = L.post.fresh + L.post.push_off[r];
    DsEntry* dst = L.post.sorted + L.post.fam_off[f];
    for (int c = threadIdx.x; c < cnt; c += blockDim.x) dst[c] = src[c];
}

// Number of entries of a sorted array that precede `key`.
__device__ inline int ds_rank(const DsEntry* a, int len, const DsEntry& key) {
    int lo = 0, hi = len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ds_less(a[mid], key)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void ds_merge_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (!S->posted || S->finished || S->pushes == 0) return;
    const int P = S->pushes, M = S->fsize;
    const DsEntry* old = L.store.f[S->fbuf] + S->fhead;
    DsEntry* out = L.store.f[S->fbuf ^ 1];
    const DsEntry* fresh = L.post.sorted;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M + P; i += gridDim.x * blockDim.x) {
        if (i < M) {
            const DsEntry e = old[i];
            out[i + ds_rank(fresh, P, e)] = e;
        } else {
            const DsEntry e = fresh[i - M];
            out[i - M + ds_rank(old, M, e)] = e;
        }
    }
}

__global__ void ds_round_end_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (!S->posted || S->finished) return;
    if (S->pushes > 0) {
        S->fbuf ^= 1;
        S->fhead = 0;
        S->fsize += S->pushes;
    }
    S->a_live = S->refilled;
    if (S->fsize == 0 && !S->refilled) {
        S->finished = 1;
        S->final_bound = S->inc;
    }
}

// B's batch (the refill) becomes A, relaxed already.
__global__ void ds_swap_kernel(DsLaunch L) {
    DsState* S = L.S;
    if (!S->posted || S->finished || !S->refilled) return;
    const int N = *L.B.count;
    const int n = L.n;
    const size_t tot = (size_t)N * n;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = g; i < tot; i += step) {
        L.A.lb[i] = L.B.lb[i];
        L.A.ub[i] = L.B.ub[i];
        L.A.x[i] = L.B.x[i];
        L.A.bv[i] = L.B.bv[i];
        L.A.b8[i] = L.B.b8[i];
        L.A.b8[tot + i] = L.B.b8[tot + i];
    }
    for (size_t i = g; i < (size_t)N; i += step) {
        L.A.eps[i] = L.B.eps[i];
        L.A.gap[i] = L.B.gap[i];
        L.A.val[i] = L.B.val[i];
        L.A.bvv[i] = L.B.bvv[i];
        L.A.cut[i] = L.B.cut[i];
        L.A.it[i] = L.B.it[i];
        L.A.st[i] = L.B.st[i];
        L.A.id[i] = L.B.id[i];
        L.A.parent[i] = L.B.parent[i];
        L.A.depth[i] = L.B.depth[i];
        L.A.riters[i] = L.B.riters[i];
        L.A.bound[i] = L.B.bound[i];
    }
    if (g == 0) {
        *L.A.count = N;
        *L.A.relax_count = 0;
    }
}

__global__ void ds_init_free_kernel(int* free_list, int C) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < C - 1; i += gridDim.x * blockDim.x)
        free_list[i] = C - 1 - i;
}

}  // namespace

int ds_launch_init_free(int* free_list, int C, hipStream_t st) {
    hipLaunchKernelGGL(ds_init_free_kernel, dim3(1024), dim3(256), 0, st, free_list, C);
    return (int)hipGetLastError();
}

int ds_launch_round_begin(const DsLaunch& L, hipStream_t st) {
    hipLaunchKernelGGL(ds_round_begin_kernel, dim3(1), dim3(1), 0, st, L);
    return (int)hipGetLastError();
}

int ds_launch_select(const DsLaunch& L, int m, hipStream_t st) {
    hipLaunchKernelGGL(ds_select_kernel, dim3(1), dim3(1), 0, st, L, m);
    hipLaunchKernelGGL(ds_stage_kernel, dim3(L.batch), dim3(256), 0, st, L, m);
    hipLaunchKernelGGL(ds_free_dropped_kernel, dim3(256), dim3(256), 0, st, L, m);
    return (int)hipGetLastError();
}

int ds_launch_post(const DsLaunch& L, hipStream_t st) {
    hipLaunchKernelGGL(ds_post_node_kernel, dim3(L.batch), dim3(64), 0, st, L);
    hipLaunchKernelGGL(ds_post_scan_kernel, dim3(1), dim3(1024), 0, st, L);
    hipLaunchKernelGGL(ds_post_emit_kernel, dim3(L.batch), dim3(256), 0, st, L);
    hipLaunchKernelGGL(ds_sort_families_kernel, dim3(1), dim3(1024), 0, st, L);
    hipLaunchKernelGGL(ds_expand_kernel, dim3(L.batch), dim3(64), 0, st, L);
    hipLaunchKernelGGL(ds_merge_kernel, dim3(1024), dim3(256), 0, st, L);
    return (int)hipGetLastError();
}

int ds_launch_round_end(const DsLaunch& L, hipStream_t st) {
    hipLaunchKernelGGL(ds_round_end_kernel, dim3(1), dim3(1), 0, st, L);
    hipLaunchKernelGGL(ds_swap_kernel, dim3(512), dim3(256), 0, st, L);
    return (int)hipGetLastError();
}

}  // namespace gb
