// Multi-instance engine: device records, launch plans and the step sequence.
//
// An Engine owns a set of instances (independent simulations, each with its
// own size, degree, gamma, BC, time step and end time) on one HIP stream.
// Kernels of one polynomial degree run over all of that degree's instances in
// one launch, through block maps built here. A step of every active instance
// is one fixed launch sequence whose data dependence goes only through the
// device records, so it can be captured once as a hipGraph and replayed.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/ppesdg.h"

#ifndef PPES_CONCURRENT
#define PPES_CONCURRENT 1
#endif
#include "ppes_types.hpp"

namespace ppes {

struct HostInst {
    ppes_config cfg;
    int px = 0, py = 0, sid = 0;     // degrees and shape id (sid_px / sid_py)
    int NX = 0, NY = 0, NN = 0, S = 0;
    double hx = 0, hy = 0;
    std::vector<double> qx_x, qw_x, Dx, qx_y, qw_y, Dy, xx, yy;
    std::vector<double *> allocs;
    size_t bytes = 0;
    int ent_offset = 0;  // the solve's ent_offset switch (drive.f90:32,280)
    int flux = 0;        // face flux (FluxKind)
};

struct Group {
    int sid = 0;  // shape id of every member
    std::vector<int> members;
    Blk *d_el = nullptr, *d_slot = nullptr, *d_face = nullptr, *d_bc = nullptr, *d_red1 = nullptr, *d_red2 = nullptr;
    Blk *d_res = nullptr;  // element map of the residual kernel (its own workgroup size), by face flux kind
    long long res_off[5] = {0, 0, 0, 0, 0};  // blocks of flux kind f: [res_off[f], res_off[f + 1])
    int *d_list = nullptr;
    int res_tb = 0;  // workgroup size of the residual launches (the residual map's elements per block follow)
    int face_tb = 0;  // workgroup size of the OEDG face launches (the face map's blocks follow)
    // tiles of the stage and limiter kernels: reduction chunks per
    // workgroup (the slot and element maps' blocks follow)
    int stage_m = 0, limit_m = 0;
    long long n_el = 0, n_slot = 0, n_face = 0, n_bc = 0, n_red1 = 0, n_res = 0;
    bool eo = false;  // some member solves with ent_offset: the stage kernels' EO variants
    // OEDG participation: some member applies it on the final stage; the
    // members that apply it on every stage (all, none, or the sub-plan
    // early_groups[early], whose maps cover only them)
    bool oedg_final = false;
    int n_every = 0, early = -1;
};

// Device buffers of one degree's block maps, kept across plan rebuilds and
// grown when needed (a batch whose membership changes every step then
// allocates nothing); the host copies stay alive while an asynchronous upload
// may read them.
struct MapPool {
    Blk *d[7] = {};
    size_t cap[7] = {};
    std::vector<Blk> h[7];
    int *dl = nullptr;
    size_t lcap = 0;
    std::vector<int> hl;
};

enum KTimerId { KT_RESIDUAL = 0, KT_STAGE = 1, KT_OEDG = 2, KT_LIMIT = 3, KT_REDUCE = 4, KT_BC = 5, KT_COUNT = 6 };

struct KTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
    double total_ms = 0.0;
    long long count = 0;
};

class Engine {
public:
    explicit Engine(int rk_stages);
    ~Engine();
    int init();                                   // stream and record storage
    int add(const ppes_config &cfg, int *id);     // allocate an instance (inactive)
    int remove(int id);
    bool valid(int id) const { return id >= 0 && id < (int)hosts.size() && hosts[id] != nullptr; }
    HostInst &host(int id) { return *hosts[id]; }
    Inst &rec(int id) { return h_recs[id]; }      // host mirror of the record

    int push();                                   // host mirror -> device
    int pull();                                   // device -> host mirror (synchronises)
    int sync();

    // launch helpers over an explicit instance list (plans are built on demand)
    int launch_startup(const std::vector<int> &ids);             // drive.f90:66-73
    int launch_step(bool all_kernels);                           // one RK step of every active instance
    // the degree groups' steps on streams of their own, forked from and
    // joined back to the engine's stream (PPES_CONCURRENT, default on; off
    // while kernels are timed)
    bool concurrent = PPES_CONCURRENT != 0;
    int run_steps(long long nsteps, bool use_graph, long long *done_steps);
    int sync_every = 1;                                          // steps queued per host synchronisation
    int arith = PPES_ARITH_FAST;                                 // residual arithmetic (PPES_ARITH_*)
    // the step's residual launches of L(u_st), st < S, store stage st + 1's
    // combined state for the stage kernel (precombine_ok); same bits either way
    bool precomb = true;
    void set_precombine(bool on)
    {
        precomb = on;
        graph_ok = false;
    }
    void set_arith(int a)
    {
        arith = a;
        graph_ok = false;
        built_key.clear();  // the residual maps depend on the mode (tuned sizes are fast-mode only)
        plan_ok = false;
    }
    // Autotuner, for each degree group of the current plan: the residual
    // (when every member uses the fast ES-HLL residual) at 64, 128 and 256
    // lanes per workgroup, the OEDG face sums at 64, 128 and 256, and the
    // stage and limiter kernels at one, two, four and eight reduction chunks
    // per workgroup (up to TILE_LANES_MAX lanes), each timed on the current
    // states and kept when fastest for that shape; tb_out[sid - 1] receives
    // the residual's choice (0: not tuned). The timing runs write only
    // scratch slots (L(u^n) into residual slot 1, the first stage into state
    // slot 1, the limiter into slot S, which the commit left holding the old
    // u^n) and the instance records are put back afterwards.
    int autotune(int reps, int *tb_out);
    int tuned_tb[NSHAPES + 1] = {};  // by shape id; 0: the default size
    int residual_tb(int sid, const std::vector<int> &members) const;
    int tuned_face_tb[NSHAPES + 1] = {};  // OEDG face workgroup size by shape id; 0: BLOCK
    int face_tb(int sid) const { return arith == PPES_ARITH_FAST && tuned_face_tb[sid] ? tuned_face_tb[sid] : BLOCK; }
    // chunks per workgroup of the stage and limiter kernels by shape id (0:
    // the 256-lane default); the ent_offset variants and the one-instance
    // maps of the reference-API calls keep the default
    int tuned_stage_m[NSHAPES + 1] = {}, tuned_limit_m[NSHAPES + 1] = {};
    int stage_mult(int sid, bool eo) const;
    int limit_mult(int sid, bool eo) const;
    // the workgroup size (lanes) a kernel class (KT_RESIDUAL, KT_STAGE,
    // KT_OEDG, KT_LIMIT) runs at for a shape id in the step plan, and a
    // manual choice of it (PPES_BAD_ARG if no variant of that size is built)
    int tile_of(int kernel, int sid) const;
    int set_tile(int kernel, int sid, int lanes);
    int launch_api(int id, int what, int slot, int flags, int ut_slot);  // single reference-API calls
    int set_flux(int id, int kind);                                      // face flux of one instance (FluxKind)

    int active_count() const;
    void invalidate_plan() { plan_ok = false; }  // re-check the active set before the next launch

    // per-kernel-class timers (HIP events on the engine's stream)
    bool timing = false;
    KTimer timers[KT_COUNT];
    void harvest(int k);

    hipStream_t stream = nullptr;
    int S;
    std::vector<HostInst *> hosts;
    Inst *h_recs = nullptr;  // pinned, capacity entries
    Inst *d_recs = nullptr;
    int cap = 0;

private:
    int ensure_capacity(int n);
    int build_plan();
    // tuned = false: the default tiles of the stage and limiter kernels
    int build_group(Group &g, const std::vector<int> &members, int sid, MapPool *pool = nullptr, bool tuned = true);
    MapPool plan_pools[NSHAPES + 1], start_pools[NSHAPES + 1];  // by shape id: the step plan's maps, the start-up's
    MapPool early_pools[NSHAPES + 1];                 // by shape id: the every-stage OEDG sub-plans' maps
    std::vector<Group> early_groups;        // every-stage OEDG members of a group, when not all of them
    // the maps that cover exactly the instances of group g that may apply
    // OEDG on stage st (the group itself or its sub-plan), or none
    const Group *oedg_group(const Group &g, int st) const
    {
        if (st == S) return g.oedg_final ? &g : nullptr;
        if (g.n_every == 0) return nullptr;
        return g.early < 0 ? &g : &early_groups[g.early];
    }
    size_t all_cap = 0;
    std::vector<int> h_all;
    long long plan_steps = 0;  // steps launched with the current plan (the graph is captured from the second)
    int capture();
    void free_group(Group &g);
    int launch_group_step(Group &g, bool any_retry, hipStream_t stream);  // one step of one group
    int ensure_group_streams(size_t n);
    std::vector<hipStream_t> gstreams;  // default priority, non-blocking
    std::vector<hipEvent_t> gevents;
    hipEvent_t fork_ev = nullptr;
    int *d_all = nullptr;
    int n_all = 0;
    std::vector<Group> groups;
    std::vector<Group> api_groups;     // one-instance maps of the reference-API calls
    std::vector<long long> built_key;  // active set of the built plan
    int built_n_all = 0;
    bool plan_ok = false;
    bool any_oedg_stage[MAXS + 1] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_ok = false;
};

enum ApiCall { API_BC = 0, API_POST = 1, API_RESID = 2, API_COMBINE = 3, API_GHOST_SAVE = 4, API_RESET = 5 };

// Side rules, nozzle count and fixed state of a PPES_BC_* kind (the BC_handle
// fillers of problems.f90:312-654) for an instance with nx columns and ratio
// of specific heats g. Returns PPES_BAD_ARG for an unknown kind or a missing
// PPES_BC_FIXED state.
int set_bc_rules(Inst &r, int kind, const double *state, int nx, double g);

// The residual's constants of an instance record (Inst::r2hx .. cface) from
// its hx, hy, g and GLL weights; called again whenever g changes.
void record_constants(Inst &r);

// DG_semidiscretize of a caller-owned device state into a caller-owned
// residual, with no allocation: the record and block map live in a fixed
// static device workspace (ppes_semidiscretize_noalloc). Synchronous.
int residual_noalloc(const ppes_config &cfg, int flux, int arith, const double *u, double *ut, double *dt_lim,
                     hipStream_t stream);

}  // namespace ppes
