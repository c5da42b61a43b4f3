// Multi-workgroup relaxation of one node (MultiBlock mode, K > 1) for dense
// quadratics over a box or a simplex, gfx950: K co-resident workgroups share a node and
// meet at its group barrier (gb_barrier.h), so one node's products stream
// the whole of Q at the bandwidth of many CUs instead of one, and n is not
// limited by LDS (multiblock.cuh, docs/multiblock.md).
//
// The relaxation is the plain form's, step for step (relax_fp64.hip), and
// its arithmetic is the one-workgroup kernel's: every product coordinate is
// the same k-ordered FMA chain (a thread computes whole rows), elementwise
// work is the same per coordinate, and every sum is taken in the canonical
// order of a workgroup of Tv threads (Tv = the one-workgroup launch's block
// size): virtual lane l accumulates coordinates l, l + Tv, ... in order, each
// virtual wave (64 consecutive global threads, one real wave) reduces with
// the same shuffle tree, and every workgroup adds the wave totals in wave
// order.  Minima and the logical ands are order-free.  The outputs therefore
// equal the one-workgroup kernel's bit for bit, whatever K.
//
// Collectives: a workgroup writes its partial (wave totals, or one value
// per workgroup) into the node's reduction slots, the group syncs, and every
// workgroup combines the slots itself, so all workgroups hold the same
// scalars and take the same branches.  The slots rotate over three buffers:
// a buffer is rewritten only two syncs after its last read.  A failed sync
// (timeout of the bounded wait, which needs a broken co-residency) turns the
// remaining collectives into no-ops so the kernel ends, and the node's
// status becomes 0xFF.
#include "gb_barrier.h"
#include "gb_device.h"
#include "gb_internal.h"

namespace gb {

namespace {

// Loads of Q per batch of a product row (two batches in flight).
constexpr int kMbRowLoads = 16;

// Global-memory view of a pointer, so loads retire in order.
typedef const __attribute__((address_space(1))) double gdouble;

struct Mb {
    const RelaxParams& P;
    const MbParams& M;
    int node, rank, gtid, G, n;
    int* flag;       // LDS word of group_sync
    double* sred;    // LDS block scratch
    int call;
    bool ok;

    __device__ bool sync() {
        if (!ok) return false;
        ok = group_sync(M.bars + node, (unsigned)M.K, M.spin, M.bst, flag);
        return ok;
    }
    __device__ double* slots() {
        double* r = M.red + ((size_t)node * 3 + (call % 3)) * kMbSlots;
        ++call;
        return r;
    }
    // Canonical sum of the virtual lanes' values (meaningful for gtid < Tv).
    __device__ double sum(double v) {
        double* r = slots();
        if (gtid < M.Tv) {
            const double w = wave_tree_sum(v);
            if ((gtid & (kWave - 1)) == 0) r[gtid / kWave] = w;
        }
        sync();
        double t = r[0];
        const int W = M.Tv / kWave;
        for (int w = 1; w < W; ++w) t = t + r[w];
        return t;
    }
    __device__ double minimum(double v) {
        double* r = slots();
        const double b = block_min(v, sred);
        if (threadIdx.x == 0) r[16 + rank] = b;
        sync();
        double t = r[16];
        for (int k = 1; k < M.K; ++k) t = fmin(t, r[16 + k]);
        return t;
    }
    __device__ bool all(bool p) {
        double* r = slots();
        int* ired = reinterpret_cast<int*>(sred);
        const bool b = block_and(p, ired);
        if (threadIdx.x == 0) r[16 + rank] = b ? 1.0 : 0.0;
        sync();
        bool t = true;
        for (int k = 0; k < M.K; ++k) t = t && r[16 + k] != 0.0;
        return t;
    }

    // Row j of Q z: the k-ordered FMA chain of the one-workgroup kernel
    // (z read from global memory instead of LDS: the same values).  A node
    // has only n such chains, so each thread streams its row in batches of
    // kMbRowLoads loads of Q and issues the next batch before it consumes
    // the current one (global loads retire in order, so the FMAs wait only
    // for their own batch).  The chain's order is unchanged.
    __device__ static void fetch(gdouble* q, int k, int n, double* out) {
#pragma unroll
        for (int u = 0; u < kMbRowLoads; ++u) out[u] = q[(size_t)(k + u) * n];
    }
    __device__ static double chain(const double* qq, gdouble* z, int k, double acc) {
#pragma unroll
        for (int u = 0; u < kMbRowLoads; ++u) acc = fma(qq[u], z[k + u], acc);
        return acc;
    }
    __device__ double row(const double* zp, int j) const {
        gdouble* q = (gdouble*)(P.Qt + j);
        gdouble* z = (gdouble*)zp;
        const int full = n / kMbRowLoads;
        double acc = 0.0, a[kMbRowLoads], b[kMbRowLoads];
        if (full > 0) fetch(q, 0, n, a);
        int t = 0;
        for (; t + 2 <= full; t += 2) {
            fetch(q, (t + 1) * kMbRowLoads, n, b);
            acc = chain(a, z, t * kMbRowLoads, acc);
            if (t + 2 < full) fetch(q, (t + 2) * kMbRowLoads, n, a);
            acc = chain(b, z, (t + 1) * kMbRowLoads, acc);
        }
        if (t < full) acc = chain(a, z, t * kMbRowLoads, acc);
        for (int k = full * kMbRowLoads; k < n; ++k) acc = fma(q[(size_t)k * n], z[k], acc);
        return acc;
    }
    // Rows go to threads wave by wave across the workgroups (row wave w to
    // workgroup w mod K), so a product of fewer than G rows still spreads
    // over every workgroup's CU.  Any row-to-thread map gives the same rows.
    __device__ int row0() const {
        return ((int)(threadIdx.x / kWave) * M.K + rank) * kWave + (int)(threadIdx.x & (kWave - 1));
    }
    // out = Q z on this group (z complete on entry); synced on return.
    __device__ void product(const double* z, double* out) {
        for (int j = row0(); j < n; j += G) out[j] = row(z, j);
        sync();
    }
    // Obj<GB_OBJ_DENSE_QUADRATIC>::grad
    __device__ void grad(const double* x, double* g) {
        for (int j = row0(); j < n; j += G) {
            const double acc = row(x, j);
            const double lin = P.lin ? P.ls * P.lin[j] : 0.0;
            g[j] = P.gs * (acc + lin);
        }
        sync();
    }
    // Obj<GB_OBJ_DENSE_QUADRATIC>::value (x complete on entry)
    __device__ double value(const double* x, double* prod) {
        product(x, prod);
        const double hs = 0.5 * P.gs, ls = P.gs * P.ls;
        double val = 0.0;
        if (gtid < M.Tv) {
            double quad = 0.0, lin = 0.0;
            for (int j = gtid; j < n; j += M.Tv) {
                quad = fma(x[j], prod[j], quad);
                if (P.lin) lin = fma(P.lin[j], x[j], lin);
            }
            val = fma(hs, quad, ls * lin);
            if (gtid == 0) val = val + P.kappa;
        }
        return sum(val);
    }
    // Obj<GB_OBJ_DENSE_QUADRATIC>::curvature (d complete on entry)
    __device__ double curvature(const double* d, double* prod) {
        product(d, prod);
        double c = 0.0;
        if (gtid < M.Tv)
            for (int j = gtid; j < n; j += M.Tv) c = fma(P.gs * d[j], prod[j], c);
        return sum(c);
    }
    // sum over j of a[j] * b[j] in the canonical order (a, b complete)
    __device__ double dot(const double* a, const double* b) {
        double acc = 0.0;
        if (gtid < M.Tv)
            for (int j = gtid; j < n; j += M.Tv) acc = fma(a[j], b[j], acc);
        return sum(acc);
    }
};


// ---------------------------------------------------------------------------
// LMOs on the group.  Each mirrors Lmo<...> of the one-workgroup kernel: the
// same per-coordinate arithmetic, sums in the canonical order, and the same
// ranks and scan, so the vertices are the same bits.
// ---------------------------------------------------------------------------
template <int LMO>
struct MbLmo;

// BoxLMO (lmo.cuh:212-321), as Lmo<GB_LMO_BOX>.
template <>
struct MbLmo<GB_LMO_BOX> {
    Mb& c;
    __device__ MbLmo(Mb& cc, unsigned char*) : c(cc) {}
    __device__ bool feasible(const double*, const double*) { return true; }
    __device__ void compute_v(const double* g, const double* lb, const double* ub, double* v) {
        for (int j = c.gtid; j < c.n; j += c.G) v[j] = (g[j] > 0.0) ? lb[j] : ub[j];
        c.sync();
    }
    // in-face vertex for the direction -g (DICG away vertex)
    __device__ void inface_neg(const double* g, const double* x, const double* lb, const double* ub, double* a) {
        const double atol = 1e-6, rtol = 1e-4;
        for (int j = c.gtid; j < c.n; j += c.G) {
            const double gj = -g[j];
            const double ut = atol + rtol * fabs(ub[j]);
            const double lt = atol + rtol * fabs(lb[j]);
            if (fabs(x[j] - ub[j]) <= ut) a[j] = ub[j];
            else if (fabs(x[j] - lb[j]) <= lt) a[j] = lb[j];
            else a[j] = (gj > 0.0) ? lb[j] : ub[j];
        }
        c.sync();
    }
    __device__ double max_step(const double* d, const double* x, const double* lb, const double* ub) {
        double r = 1.0;
        for (int j = c.gtid; j < c.n; j += c.G) {
            const double dj = d[j];
            if (dj < 0.0) r = fmin(r, (ub[j] - x[j]) / -dj);
            else if (dj > 0.0) r = fmin(r, (x[j] - lb[j]) / dj);
        }
        return fmax(0.0, fmin(c.minimum(r), 1.0));
    }
    __device__ bool rounded_feasible(const double*) { return true; }
};

// Per-node vectors of the simplex LMO (S = next_pow2(n) ranks).
struct MbSimplexWork {
    int* rank;        // [n] rank of an eligible coordinate by (key, index)
    int* sidx;        // [S] coordinate at each rank (-1: none)
    double* scap;     // [S] capacity at each rank
    double* spre;     // [S] exclusive prefix of the capacities
    uint8_t* fixed;   // [n] coordinate fixed on the current face
};

__host__ __device__ inline size_t mb_up256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of the vectors in a node's region: rank, sidx, scap, spre,
// fixed, and the total.
__host__ __device__ inline void mb_simplex_offsets(int n, int S, size_t off[6]) {
    off[0] = 0;
    off[1] = off[0] + mb_up256((size_t)n * sizeof(int));
    off[2] = off[1] + mb_up256((size_t)S * sizeof(int));
    off[3] = off[2] + mb_up256((size_t)S * sizeof(double));
    off[4] = off[3] + mb_up256((size_t)S * sizeof(double));
    off[5] = off[4] + mb_up256((size_t)n);
}

__device__ inline MbSimplexWork mb_simplex_layout(unsigned char* base, int n, int S) {
    size_t off[6];
    mb_simplex_offsets(n, S, off);
    MbSimplexWork w;
    w.rank = reinterpret_cast<int*>(base + off[0]);
    w.sidx = reinterpret_cast<int*>(base + off[1]);
    w.scap = reinterpret_cast<double*>(base + off[2]);
    w.spre = reinterpret_cast<double*>(base + off[3]);
    w.fixed = base + off[4];
    return w;
}

// Coordinates of one rank tile held in LDS.
constexpr int kMbRankTile = 1024;

// UnitSimplexLMO / EqualitySimplexLMO (lmo.cuh:335-912), as
// Lmo<GB_LMO_SIMPLEX>: the greedy allocation ranks the eligible coordinates
// by (key, index) -- each rank is an integer count of smaller keys, taken
// over LDS tiles of the keys by every workgroup of the group and added with
// integer atomics (exact in any order) -- scatters the capacities by rank,
// scans them in the one-workgroup kernel's chunked order on one wave, and
// fills every coordinate from its prefix mass.
template <>
struct MbLmo<GB_LMO_SIMPLEX> {
    Mb& c;
    MbSimplexWork w;
    int S;
    __device__ MbLmo(Mb& cc, unsigned char* base) : c(cc) {
        S = 1;
        while (S < c.n) S <<= 1;
        w = mb_simplex_layout(base, c.n, S);
    }
    __device__ bool elig(const double* g, double sgn, bool use_fixed, bool force_full, int j) const {
        return (!use_fixed || !w.fixed[j]) && (force_full || sgn * g[j] <= 0.0);
    }
    // out holds its base values on entry (complete); synced on return.
    __device__ void greedy(const double* g, double sgn, const double* ub, double* out, bool use_fixed,
                           double remaining, bool force_full, bool floor_int) {
        const RelaxParams& P = c.P;
        const int n = c.n, G = c.G, gtid = c.gtid;
        const uint8_t* is_int = P.is_int;
        __shared__ double tkey[kMbRankTile];
        __shared__ uint8_t tel[kMbRankTile];
        for (int p = gtid; p < S; p += G) { w.sidx[p] = -1; w.scap[p] = 0.0; }
        for (int j = gtid; j < n; j += G) w.rank[j] = 0;
        c.sync();
        // ranks: job (j block of kMbThreads, i tile of kMbRankTile) round-robin over the workgroups
        const int jb = (n + kMbThreads - 1) / kMbThreads;
        const int ib = (n + kMbRankTile - 1) / kMbRankTile;
        for (int job = c.rank; job < jb * ib; job += c.M.K) {
            const int j0 = (job / ib) * kMbThreads, i0 = (job % ib) * kMbRankTile;
            const int len = min(kMbRankTile, n - i0);
            for (int t = threadIdx.x; t < len; t += blockDim.x) {
                tkey[t] = sgn * g[i0 + t];
                tel[t] = elig(g, sgn, use_fixed, force_full, i0 + t) ? 1 : 0;
            }
            __syncthreads();
            const int j = j0 + (int)threadIdx.x;
            if (j < n && elig(g, sgn, use_fixed, force_full, j)) {
                const double kj = sgn * g[j];
                int cnt = 0;
                for (int t = 0; t < len; ++t) {
                    const double ki = tkey[t];
                    cnt += (tel[t] && (ki < kj || (ki == kj && i0 + t < j))) ? 1 : 0;
                }
                if (cnt) atomicAdd(&w.rank[j], cnt);
            }
            __syncthreads();
        }
        c.sync();
        for (int j = gtid; j < n; j += G) {
            if (!elig(g, sgn, use_fixed, force_full, j)) continue;
            const int r = w.rank[j];
            w.sidx[r] = j;
            w.scap[r] = is_int[j] ? fmax(ub[j] - out[j], 0.0) : remaining;
        }
        c.sync();
        if (gtid < kWave) {
            const int l = gtid;
            const int C = (S + kWave - 1) / kWave;
            double t = 0.0;
            for (int i = 0; i < C; ++i) {
                const int p = l * C + i;
                if (p < S) t = t + w.scap[p];
            }
            double inc = t;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const double o = __shfl_up(inc, off, kWave);
                if (l >= off) inc = inc + o;
            }
            double run = __shfl_up(inc, 1, kWave);
            if (l == 0) run = 0.0;
            for (int i = 0; i < C; ++i) {
                const int p = l * C + i;
                if (p >= S) break;
                w.spre[p] = run;
                run = run + w.scap[p];
            }
        }
        c.sync();
        const bool eq = P.equality != 0;
        for (int p = gtid; p < S; p += G) {
            const int j = w.sidx[p];
            if (j < 0) continue;
            double add = fmax(remaining - w.spre[p], 0.0);
            const double cp = is_int[j] ? fmax(ub[j] - out[j], 0.0) : remaining;
            add = fmin(add, cp);
            if (floor_int && is_int[j]) add = floor(add + (eq ? 1e-9 : 0.0));
            out[j] = out[j] + add;
        }
        c.sync();
    }
    __device__ bool feasible(const double* lb, const double* ub) {
        const RelaxParams& P = c.P;
        bool valid = true, cont = false;
        double mand = 0.0, capsum = 0.0;
        for (int j = c.gtid; j < c.n; j += c.G) {
            const bool integer = P.flags ? P.flags[j] != 0 : P.is_int[j] != 0;
            const double lo = integer ? ceil(fmax(0.0, lb[j])) : fmax(0.0, lb[j]);
            const double hi = integer ? floor(ub[j]) : ub[j];
            valid = valid && lo <= hi;
            cont = cont || !integer;
        }
        if (c.gtid < c.M.Tv)
            for (int j = c.gtid; j < c.n; j += c.M.Tv) {
                const bool integer = P.flags ? P.flags[j] != 0 : P.is_int[j] != 0;
                mand = mand + (integer ? ceil(fmax(0.0, lb[j])) : fmax(0.0, lb[j]));
                capsum = capsum + (integer ? floor(ub[j]) : ub[j]);
            }
        const bool iv = c.all(valid);
        const double ms = c.sum(mand);
        bool feas = iv && ms <= P.rhs + 1e-9;
        if (P.equality) {
            const double cs = c.sum(capsum);
            const bool hc = !c.all(!cont);
            feas = feas && cs >= P.rhs - 1e-9 && (hc || fabs(P.rhs - round(P.rhs)) <= 1e-9);
        }
        return feas;
    }
    __device__ void compute_v(const double* g, const double* lb, const double* ub, double* v) {
        const RelaxParams& P = c.P;
        for (int j = c.gtid; j < c.n; j += c.G) v[j] = P.is_int[j] ? lb[j] : 0.0;
        double mass = 0.0;
        if (c.gtid < c.M.Tv)
            for (int j = c.gtid; j < c.n; j += c.M.Tv) mass = mass + (P.is_int[j] ? lb[j] : 0.0);
        const double total = c.sum(mass);  // its sync also publishes v
        const bool eq = P.equality != 0;
        greedy(g, 1.0, ub, v, false, P.rhs - total, eq, eq);
    }
    __device__ void inface_neg(const double* g, const double* x, const double* lb, const double* ub, double* a) {
        const RelaxParams& P = c.P;
        const double atol = 1e-6, rtol = 1e-4;
        auto face = [&](int j, uint8_t& fx) {
            fx = 0;
            const bool ii = P.is_int[j] != 0;
            double aj = ii ? lb[j] : 0.0;
            const double lower = ii ? lb[j] : 0.0;
            const double ltol = atol + rtol * fmax(fabs(x[j]), fabs(lower));
            if (fabs(x[j] - lower) <= ltol) { fx = 1; aj = lower; }
            else if (ii) {
                const double utol = atol + rtol * fmax(fabs(x[j]), fabs(ub[j]));
                if (fabs(x[j] - ub[j]) <= utol) { fx = 1; aj = ub[j]; }
            }
            return aj;
        };
        for (int j = c.gtid; j < c.n; j += c.G) {
            uint8_t fx;
            a[j] = face(j, fx);
            w.fixed[j] = fx;
        }
        double mass = 0.0, xmass = 0.0;
        if (c.gtid < c.M.Tv)
            for (int j = c.gtid; j < c.n; j += c.M.Tv) {
                uint8_t fx;
                mass = mass + face(j, fx);
                xmass = xmass + x[j];
            }
        const double m = c.sum(mass);  // its sync also publishes a and fixed
        const double xm = c.sum(xmass);
        const bool force = P.equality || fabs(xm - P.rhs) <= atol + rtol * fmax(fabs(xm), fabs(P.rhs));
        greedy(g, -1.0, ub, a, true, P.rhs - m, force, true);
    }
    __device__ double max_step(const double* d, const double* x, const double* lb, const double* ub) {
        const RelaxParams& P = c.P;
        const double tol = 1e-6;
        double r = 1.0;
        for (int j = c.gtid; j < c.n; j += c.G) {
            const double dj = d[j];
            const bool ii = P.is_int[j] != 0;
            if (dj > tol) r = fmin(r, (x[j] - (ii ? lb[j] : 0.0)) / dj);
            else if (dj < -tol) r = fmin(r, ((ii ? ub[j] : P.rhs) - x[j]) / -dj);
        }
        const double coord = c.minimum(r);
        double sx = 0.0, sd = 0.0;
        if (c.gtid < c.M.Tv)
            for (int j = c.gtid; j < c.n; j += c.M.Tv) sx = sx + x[j];
        const double tx = c.sum(sx);
        if (c.gtid < c.M.Tv)
            for (int j = c.gtid; j < c.n; j += c.M.Tv) sd = sd + d[j];
        const double td = c.sum(sd);
        double agg;
        if (P.equality && fabs(td) > 1e-9) agg = 0.0;
        else agg = td < 0.0 ? fmin(1.0, (tx - P.rhs) / td) : 1.0;
        return fmax(0.0, fmin(coord, agg));
    }
    __device__ bool rounded_feasible(const double* x) {
        const RelaxParams& P = c.P;
        bool valid = true;
        for (int j = c.gtid; j < c.n; j += c.G) valid = valid && x[j] >= -P.ctol;
        double sum = 0.0;
        if (c.gtid < c.M.Tv)
            for (int j = c.gtid; j < c.n; j += c.M.Tv) sum = sum + x[j];
        const bool nn = c.all(valid);
        const double t = c.sum(sum);
        return nn && (P.equality ? fabs(t - P.rhs) <= P.ctol : t <= P.rhs + P.ctol);
    }
};

__device__ inline double mb_quad_exact(double gap, double curvature, double maximum) {
    if (maximum <= 0.0) return 0.0;
    if (curvature > 0.0) {
        if (gap <= 0.0) return 0.0;
        const double step = gap / curvature;
        return step < maximum ? step : maximum;
    }
    return -gap + 0.5 * curvature * maximum < 0.0 ? maximum : 0.0;
}

template <int LMO>
__global__ void __launch_bounds__(kMbThreads) relax_multiblock_kernel(RelaxParams P, MbParams M) {
    __shared__ int flag;
    __shared__ double sred[kMaxWaves];
    const int local = blockIdx.x / M.K;
    const int node = M.node0 + local;
    const int rank = blockIdx.x % M.K;
    const int n = P.n;
    Mb c{P, M, local, rank, rank * (int)blockDim.x + (int)threadIdx.x, M.K * (int)blockDim.x, n, &flag, sred, 0, true};
    if (node >= P.N) return;  // whole groups only (the launch covers complete groups)
    const int gtid = c.gtid, G = c.G;
    const double eps = P.eps[node];
    const int initial = P.iters[node];
    const bool supplied = (P.status[node] & GB_NODE_INITIAL_POINT_SUPPLIED) != 0;
    double* x = P.x + (size_t)node * n;
    const double* lb = P.lb + (size_t)node * n;
    const double* ub = P.ub + (size_t)node * n;
    double* best_v = P.best_v + (size_t)node * n;
    double* g = M.work + (size_t)local * 6 * n;
    double* v = g + n;
    double* aw = v + n;
    double* tx = aw + n;
    double* tg = tx + n;
    double* prod = tg + n;
    MbLmo<LMO> L(c, M.lwork + (size_t)local * M.lwork_stride);

    // node feasibility before any objective access
    bool valid = true;
    for (int j = gtid; j < n; j += G)
        if (isnan(lb[j]) || isnan(ub[j]) || lb[j] > ub[j] || (P.flags && P.flags[j] && ceil(lb[j]) > floor(ub[j])))
            valid = false;
    bool feasible = c.all(valid);
    if (feasible) feasible = L.feasible(lb, ub);
    if (!feasible) {
        if (rank == 0 && threadIdx.x == 0) {
            P.status[node] = c.ok ? GB_NODE_INFEASIBLE : 0xFF;
            P.value[node] = __builtin_inf();
            P.gap[node] = 0.0;
            P.best_val[node] = kNone;
            P.iters[node] = initial;
        }
        return;
    }

    // drs_integer_value
    auto integer_value = [&](const double* vv) {
        bool ok = true;
        for (int j = gtid; j < n; j += G) {
            const double vj = vv[j];
            if (!isfinite(vj) || vj < lb[j] - P.ctol || vj > ub[j] + P.ctol
                    || (P.flags && P.flags[j] && vj != round(vj)))
                ok = false;
        }
        if (!c.all(ok)) return kNone;
        const double val = c.value(vv, prod);
        return isfinite(val) ? val : kNone;
    };
    double best_s = kNone;
    auto track = [&](const double* vert) {
        const double fv = integer_value(vert);
        if (fv < best_s) best_s = fv;
        if (fv < 1e299 && fv == best_s)
            for (int j = gtid; j < n; j += G) best_v[j] = vert[j];
    };
    auto compute_v = [&](const double* gg, double* out) { L.compute_v(gg, lb, ub, out); };

    // SecantLS::search (algorithms.cuh:121-258), as the one-workgroup kernel
    auto secant = [&](double gap, double max_gamma, const double* dir) {
        const double tol = P.secant_tol;
        const double f_base = c.value(x, prod);
        double lo = 0.0, hi = max_gamma, best_f = f_base, best_gamma = 0.0;
        double phi_lo = -gap, phi_hi = 0.0, gamma = 0.0;
        bool done = phi_lo >= -tol;
        const double f_abs = f_base >= 0.0 ? f_base : -f_base;
        const double f_allow = tol * (1.0 + f_abs);
        auto trial = [&](double gm, double& phi, double& f) {
            for (int j = gtid; j < n; j += G) tx[j] = fma(gm, dir[j], x[j]);
            c.sync();
            c.grad(tx, tg);
            phi = c.dot(tg, dir);
            f = c.value(tx, prod);
        };
        double f_hi;
        trial(max_gamma, phi_hi, f_hi);
        if (f_hi < best_f || (best_gamma == 0.0 && f_hi <= f_base + f_allow)) {
            best_f = f_hi;
            best_gamma = max_gamma;
        }
        if (!done && phi_hi <= tol && f_hi <= f_base + f_allow) { gamma = max_gamma; done = true; }
        for (int it = 0; it < P.secant_iters; ++it) {
            if (done) break;
            const double denom = phi_hi - phi_lo;
            double gm = (denom > P.secant_denom || denom < -P.secant_denom)
                ? hi - phi_hi * (hi - lo) / denom : 0.5 * (lo + hi);
            if (!(gm > lo && gm < hi)) gm = 0.5 * (lo + hi);
            gamma = gm;
            double phi, ft;
            trial(gamma, phi, ft);
            if (ft < best_f || (best_gamma == 0.0 && ft <= f_base + f_allow)) { best_f = ft; best_gamma = gamma; }
            if (phi < 0.0) { lo = gamma; phi_lo = phi; } else { hi = gamma; phi_hi = phi; }
            if ((phi <= tol && phi >= -tol) || (hi - lo) <= tol) { gamma = 0.5 * (lo + hi); done = true; }
        }
        if (best_gamma > 0.0 && best_f <= f_base + f_allow) gamma = best_gamma;
        else if (!done) gamma = 0.0;
        return gamma;
    };

    if (initial == 0 && !supplied) {
        for (int j = gtid; j < n; j += G) g[j] = -1.0 - (double)j / (n + 1.0);
        compute_v(g, v);
        for (int j = gtid; j < n; j += G) x[j] = v[j];
    }
    c.sync();

    const bool secant_like = P.ls_kind == GB_LS_SECANT || P.ls_kind == GB_LS_QUADRATIC_EXACT;
    int iterations = initial;
    for (int iter = 0; iter < P.max_iters; ++iter) {
        iterations = initial + iter + 1;
        c.grad(x, g);
        compute_v(g, v);
        track(v);
        for (int j = gtid; j < n; j += G) tx[j] = x[j] - v[j];
        c.sync();
        const double gap = c.dot(g, tx);
        if (gap >= 0.0 && gap < eps) break;
        if (!P.dicg) {
            if (secant_like) {
                for (int j = gtid; j < n; j += G) v[j] = v[j] - x[j];
                c.sync();
                double gamma;
                if (P.ls_kind == GB_LS_SECANT) gamma = secant(gap, 1.0, v);
                else gamma = mb_quad_exact(gap, c.curvature(v, prod), 1.0);
                for (int j = gtid; j < n; j += G) x[j] = fma(gamma, v[j], x[j]);
            } else {
                const double gamma = 2.0 / (double)(initial + iter + 2);
                for (int j = gtid; j < n; j += G) x[j] = fma(gamma, v[j] - x[j], x[j]);
            }
            c.sync();
            continue;
        }
        // the in-face (away) vertex of -g
        L.inface_neg(g, x, lb, ub, aw);
        track(aw);
        for (int j = gtid; j < n; j += G) aw[j] = aw[j] - v[j];
        c.sync();
        const double gmax = L.max_step(aw, x, lb, ub);
        if (secant_like) {
            for (int j = gtid; j < n; j += G) {
                aw[j] = -aw[j];
                tg[j] = -g[j];
            }
            c.sync();
            const double pgap = c.dot(tg, aw);
            double gamma;
            if (P.ls_kind == GB_LS_SECANT) gamma = secant(pgap, gmax, aw);
            else gamma = mb_quad_exact(pgap, c.curvature(aw, prod), gmax);
            for (int j = gtid; j < n; j += G) x[j] = fma(gamma, aw[j], x[j]);
        } else {
            const double ol = 2.0 / (double)(initial + iter + 2);
            double gamma;
            if (P.ls_kind == GB_LS_SAFEGUARDED_SCALED_OPEN_LOOP) gamma = gmax <= ol ? gmax : gmax * ol;
            else if (P.ls_kind == GB_LS_SCALED_OPEN_LOOP) gamma = gmax * ol;
            else gamma = fmin(gmax, ol);
            for (int j = gtid; j < n; j += G) x[j] = fma(-gamma, aw[j], x[j]);
        }
        c.sync();
    }

    // final certificate at the returned iterate
    c.grad(x, g);
    compute_v(g, v);
    track(v);
    for (int j = gtid; j < n; j += G) tx[j] = x[j] - v[j];
    c.sync();
    const double fgap = c.dot(g, tx);
    const double fval = c.value(x, prod);
    uint8_t st = (fgap >= 0.0 && fgap < eps) ? 1 : 0;
    if (P.flags) {
        for (int j = gtid; j < n; j += G) {
            double cc = x[j];
            if (P.flags[j]) cc = P.heuristic == GB_HEUR_FLOOR_OFFSET ? floor(x[j]) + P.heuristic_offset : round(x[j]);
            v[j] = cc;
        }
        c.sync();
        if (L.rounded_feasible(v)) {
            const double cv = integer_value(v);
            if (cv < best_s) {
                for (int j = gtid; j < n; j += G) best_v[j] = v[j];
                best_s = cv;
            }
        }
    }
    bool integral = true;
    for (int j = gtid; j < n; j += G)
        if (!isfinite(x[j]) || (P.flags && P.flags[j] && fabs(x[j] - round(x[j])) > P.ctol)) integral = false;
    const bool all_int = c.all(integral);
    if (all_int) st = st ? GB_NODE_INTEGRAL_CONVERGED : GB_NODE_INTEGRAL_UNCONVERGED;
    if (rank == 0 && threadIdx.x == 0) {
        P.value[node] = fval;
        P.gap[node] = fgap;
        P.status[node] = c.ok ? st : 0xFF;
        P.iters[node] = iterations;
        P.best_val[node] = best_s;
    }
}

__global__ void mb_clear_kernel(GroupBarrier* bars, int count, BarrierStatus* bst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        bars[i].arrived = 0u;
        bars[i].generation = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        bst->abort = 0u;
        bst->timeouts = 0ull;
    }
}

}  // namespace

int launch_multiblock_clear(GroupBarrier* bars, int count, BarrierStatus* bst, void* stream) {
    hipLaunchKernelGGL(mb_clear_kernel, dim3(64), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), bars, count,
                       bst);
    return (int)hipGetLastError();
}

// Workgroups of the multi-workgroup kernel that are resident on the current
// device at once.
size_t multiblock_lmo_bytes(int lmo_kind, int n) {
    if (lmo_kind != GB_LMO_SIMPLEX) return 0;
    size_t off[6];
    mb_simplex_offsets(n, next_pow2_host(n), off);
    return off[5];
}

int multiblock_capacity(int lmo_kind) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    const void* k = lmo_kind == GB_LMO_SIMPLEX ? reinterpret_cast<const void*>(relax_multiblock_kernel<GB_LMO_SIMPLEX>)
                                               : reinterpret_cast<const void*>(relax_multiblock_kernel<GB_LMO_BOX>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kMbThreads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}

int launch_relax_multiblock(const RelaxParams& p, const MbParams& m, int nodes, void* stream) {
    RelaxParams pp = p;
    MbParams mm = m;
    void* args[] = {&pp, &mm};
    const void* k = p.lmo_kind == GB_LMO_SIMPLEX ? reinterpret_cast<const void*>(relax_multiblock_kernel<GB_LMO_SIMPLEX>)
                                                 : reinterpret_cast<const void*>(relax_multiblock_kernel<GB_LMO_BOX>);
    // A plain launch: the caller sizes it to the resident capacity, and the
    // group barrier's bounded waits end a group that is not resident
    // together (status 0xFF) instead of hanging.  hipLaunchCooperativeKernel
    // was refused (hipErrorIllegalState) in a process that had run many
    // searches on streams with hardware queues of their own.
    const hipError_t e = hipLaunchKernel(k, dim3(nodes * m.K), dim3(kMbThreads), args, 0,
                                         reinterpret_cast<hipStream_t>(stream));
    return (int)e;
}

}  // namespace gb
