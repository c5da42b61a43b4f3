// Device-resident branch-and-bound search (frontier.hip kernels, dsearch.cpp
// host side): the frontier, the incumbent, pruning and branching live on the
// GPU, and every round of the search is one replay of a captured HIP graph.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gb {

// A frontier entry: the key (bound, id) orders the best-first queue exactly
// as the host driver's heap (ids are unique), slot holds the node's data.
struct DsEntry {
    double bound;
    int id;
    int slot;
};

enum DsKind { kDsInfeasible = 0, kDsBranch = 1, kDsIntegral = 2, kDsRequeue = 3 };

// Search state, one per search, in device memory.
struct DsState {
    double inc;            // incumbent value (+inf: none)
    double final_bound;    // dual bound once the search has stopped
    long long fw_iters;
    long long done_nodes;
    long long max_nodes;   // <= 0: no cap
    long long rounds;
    int next_id;
    int finished;          // the search has stopped (any reason)
    int tol_hit;           // stopped on the gap tolerance
    int limit_hit;         // stopped on the node cap or a host stop request
    int overflow;          // stopped: node slots or frontier capacity exhausted
    int stop_request;      // set by the host (time limit): stop at the next round
    int fbuf, fhead, fsize;  // frontier buffer, first entry, entries (stale ones included)
    int free_top;          // free node slots on the free list
    int a_live;            // batch A is relaxed and waits for its post-processing
    int refilled;          // batch B was filled from the frontier in this round
    int inc_row;           // row of A whose best vertex became the incumbent (-1)
    int posted;            // batch A was post-processed in this round
    int pushes;            // frontier entries this round's post-processing pushes
    int families;          // rows of A that push entries
    int plan_first[2], plan_take[2], plan_drop[2], plan_free_base[2];
};

// Node data by slot.
struct DsStore {
    uint8_t* b8;        // [C][2n]: lower then upper bounds as bytes
    double* x0;         // [C][n]: start point (own iterate of a requeued node,
                        //          the parent's iterate clipped to the box of a child)
    int* id;
    int* parent;
    int* depth;
    int* riters;        // iterations already done (continuations)
    uint8_t* warm;      // x0 holds a start point
    int* free_list;     // [C] stack of free slots
    DsEntry* f[2];      // frontier buffers, [C] entries each
    int capacity;
};

// One node batch of up to B rows.
struct DsBatch {
    int* count;         // device: rows in the batch
    int* relax_count;   // device: rows the relax launches of this round take (0: relaxed already)
    double *lb, *ub, *x, *eps, *gap, *val, *bv, *bvv, *cut;  // [B][n] and [B]
    int* it;
    uint8_t* st;
    int *id, *parent, *depth, *riters;  // node records [B]
    double* bound;      // bound the node was launched with [B]
    uint8_t* b8;        // [B][2n] bounds as bytes
};

// Post-processing of batch A, per row and per round.
struct DsPost {
    double* nb;         // certified bound
    double* cand;       // best vertex value offered as incumbent (+inf: none)
    int* kind;          // DsKind
    int* nvars;         // branching variables
    int* vars;          // [B][kDsMaxK]
    int* pushes;        // entries the row pushes (0: pruned, nothing, or after the stop)
    int* push_off;      // first of its entries in `fresh`
    int* id_base;       // id of its first child
    int* slot_base;     // free-list position of its first slot
    int* fam_row;       // rows that push, sorted by their (bound, first id)
    int* fam_off;       // output offset of each sorted family
    DsEntry* fresh;     // [16 B] entries in processing order
    DsEntry* sorted;    // [16 B] the same, sorted by (bound, id)
    double* inc_x;      // [n] incumbent
};

constexpr int kDsMaxK = 3;       // branching variables per node (2^3 children within the 8 reserved ids)
constexpr int kDsMaxBatch = 1024;

struct DsLaunch {
    DsState* S;
    DsStore store;
    DsBatch A, B;
    DsPost post;
    int n;
    int batch;                 // B: rows per batch
    const int* iv;             // integer variables, problem order
    int n_iv;
    const double* eps_table;   // node FW tolerance by depth [n + 2]
    double int_tol;
    double abs_tol, rel_tol;
    int rule;                  // GB_BRANCH_POW2 or GB_BRANCH_MOST_FRACTIONAL
    int branch_k;
    int child_warm;            // children start at the parent's clipped iterate (a box);
                               // 0: at the LMO vertex (a simplex, where clipping may break sum x <= rhs)
};

// Kernels of one round, in order; relax launches of A and B go between.
int ds_launch_round_begin(const DsLaunch& L, hipStream_t st);
int ds_launch_select(const DsLaunch& L, int refill, hipStream_t st);
int ds_launch_post(const DsLaunch& L, hipStream_t st);
int ds_launch_round_end(const DsLaunch& L, hipStream_t st);
// free_list[i] = C - 1 - i for i < C - 1 (slot 0 is the root's).
int ds_launch_init_free(int* free_list, int C, hipStream_t st);

}  // namespace gb
