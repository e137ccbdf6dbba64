// plan.hip -- the batch time loop, step by step or as replayed hipGraphs.
//
// A plan keeps a two-int device cursor {slot, seg_base}.  A graph of G steps
// is captured once per (batch contents, mode, direction, active counts): its
// step nodes carry slot offsets 0, d, .., (G-1)d (d = direction) and read the
// base slot and the dWaveOp store base from the cursor; its last node adds
// G*d to the cursor.  Replaying it n times runs n*G consecutive steps with one
// host call each.  Before the first graph chunk of a run (and after steps
// launched one by one) a one-thread kernel sets the cursor.
#include "common.h"

#include <cstring>
#include <vector>

struct pysit_amd_plan {
    int device = 0;
    int graph_steps = 0;
    int max_graphs = 0;
    int *cursor = nullptr;      // device int[2]
    hipStream_t cap = nullptr;  // capture stream (the caller's may be the null stream)
    struct Entry {
        std::vector<unsigned char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    std::vector<Entry> graphs;
    unsigned long long tick = 0;
    long long n_launch = 0, n_capture = 0, n_direct = 0;
};

namespace pa {

__global__ void k_cursor_set(int *cursor, int slot, int seg_base)
{
    if (threadIdx.x == 0) {
        cursor[0] = slot;
        cursor[1] = seg_base;
    }
}

__global__ void k_cursor_add(int *cursor, int d)
{
    if (threadIdx.x == 0) cursor[0] += d;
}

static void free_entry(pysit_amd_plan::Entry &e)
{
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
    e.exec = nullptr;
    e.graph = nullptr;
}

static std::vector<unsigned char> graph_key(const pysit_amd_batch &B, const LoopSpec &L, int nt, int nr, int G,
                                            int phase)
{
    const int tail[7] = {L.mode, L.direction, nt, nr, L.window, G, phase};
    std::vector<unsigned char> k(sizeof(B) + sizeof(tail));
    std::memcpy(k.data(), &B, sizeof(B));
    std::memcpy(k.data() + sizeof(B), tail, sizeof(tail));
    return k;
}

// The slot of local step i of a chunk, modulo the window, whose first slot
// is `phase` modulo the window; whether a window ends after it.
static bool chunk_window_end(int phase, int i, int direction, int w)
{
    const int j = ((phase + direction * i) % w + w) % w;
    return window_end(j, direction, w);
}

// The cached graph for `key`, captured on first use.
static int graph_for(pysit_amd_plan *plan, const std::vector<unsigned char> &key, const LoopSpec &L, StepSink &S,
                     int nt, int nr, int phase, hipStream_t st, hipGraphExec_t *out)
{
    for (auto &e : plan->graphs) {
        if (e.key == key) {
            e.used = ++plan->tick;
            *out = e.exec;
            return PYSIT_AMD_OK;
        }
    }
    if ((int)plan->graphs.size() >= plan->max_graphs) {
        // evict the least recently used; a launch of it may still be queued
        size_t v = 0;
        for (size_t i = 1; i < plan->graphs.size(); ++i)
            if (plan->graphs[i].used < plan->graphs[v].used) v = i;
        PA_HIP(hipStreamSynchronize(st));
        free_entry(plan->graphs[v]);
        plan->graphs.erase(plan->graphs.begin() + v);
    }
    const int G = plan->graph_steps;
    PA_HIP(hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal));
    int rc = PYSIT_AMD_OK;
    for (int i = 0; i < G && !rc; ++i) {
        rc = S.step(L.direction * i, 0, nt, nr, plan->cursor, plan->cap);
        if (!rc && L.window && chunk_window_end(phase, i, L.direction, L.window)) rc = S.flush(plan->cap);
    }
    if (!rc) {
        hipLaunchKernelGGL(k_cursor_add, dim3(1), dim3(64), 0, plan->cap, plan->cursor, L.direction * G);
        if (hipGetLastError() != hipSuccess) {
            set_error("cursor kernel launch failed during graph capture");
            rc = PYSIT_AMD_EDEVICE;
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(plan->cap, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        return PYSIT_AMD_EDEVICE;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        return PYSIT_AMD_EDEVICE;
    }
    pysit_amd_plan::Entry e;
    e.key = key;
    e.graph = g;
    e.exec = x;
    e.used = ++plan->tick;
    plan->graphs.push_back(std::move(e));
    ++plan->n_capture;
    *out = x;
    return PYSIT_AMD_OK;
}

int plan_device(const pysit_amd_plan *plan) { return plan->device; }

int time_loop(pysit_amd_plan *plan, const pysit_amd_batch &B, const LoopSpec &L, StepSink &S, hipStream_t st)
{
    const int n = L.slot_end - L.slot_begin;
    const int G = plan ? plan->graph_steps : 0;
    auto slot_of = [&](int i) { return L.direction > 0 ? L.slot_begin + i : L.slot_end - 1 - i; };
    auto tiles_at = [&](int j) { return L.active_tiles ? L.active_tiles[j] : B.n_tiles; };
    auto recs_at = [&](int j) { return L.active_recs ? L.active_recs[j] : B.n_rec; };
    // chunk [i0, i0 + G) can be a graph: whole, window-aligned, constant counts
    auto uniform = [&](int i0, int nt, int nr) {
        if (G <= 0 || i0 + G > n || (L.window && G % L.window)) return false;
        for (int i = i0; i < i0 + G; ++i)
            if (tiles_at(slot_of(i)) != nt || recs_at(slot_of(i)) != nr) return false;
        return true;
    };
    bool cursor_known = false;
    int cursor_slot = 0;
    int i0 = 0;
    while (i0 < n) {
        const int j0 = slot_of(i0);
        const int nt = tiles_at(j0), nr = recs_at(j0);
        bool as_graph = uniform(i0, nt, nr);
        // window phase of the chunk's first slot: graphs of chunks at another
        // phase flush after other local steps
        const int phase = L.window ? j0 % L.window : 0;
        std::vector<unsigned char> key;
        if (as_graph) {
            key = graph_key(B, L, nt, nr, G, phase);
            bool cached = false;
            for (auto &e : plan->graphs) cached = cached || e.key == key;
            // a new graph pays for its capture only if it is replayed
            if (!cached) as_graph = uniform(i0 + G, nt, nr);
        }
        if (!as_graph) {
            const int len = G > 0 ? (n - i0 < G ? n - i0 : G) : n - i0;
            for (int i = i0; i < i0 + len; ++i) {
                const int j = slot_of(i);
                int rc = S.step(j, L.seg_base, tiles_at(j), recs_at(j), nullptr, st);
                if (!rc && L.window && window_end(j, L.direction, L.window)) rc = S.flush(st);
                if (rc) return rc;
            }
            if (plan) plan->n_direct += len;
            i0 += len;
            continue;
        }
        hipGraphExec_t x = nullptr;
        int rc = graph_for(plan, key, L, S, nt, nr, phase, st, &x);
        if (rc) return rc;
        if (!cursor_known || cursor_slot != j0) {
            hipLaunchKernelGGL(k_cursor_set, dim3(1), dim3(64), 0, st, plan->cursor, j0, L.seg_base);
            PA_HIP(hipGetLastError());
        }
        PA_HIP(hipGraphLaunch(x, st));
        ++plan->n_launch;
        cursor_known = true;
        cursor_slot = j0 + L.direction * G;
        i0 += G;
    }
    return PYSIT_AMD_OK;
}

}  // namespace pa

using namespace pa;

extern "C" {

int pysit_amd_plan_create(int graph_steps, int max_graphs, pysit_amd_plan **out)
{
    clear_error();
    PA_CHECK(out != nullptr, "null output pointer");
    *out = nullptr;
    PA_CHECK(graph_steps >= 1 && graph_steps <= 4096, "graph_steps must be in 1..4096");
    PA_CHECK(max_graphs >= 1 && max_graphs <= 1024, "max_graphs must be in 1..1024");
    pysit_amd_plan *p = new pysit_amd_plan();
    p->graph_steps = graph_steps;
    p->max_graphs = max_graphs;
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipMalloc((void **)&p->cursor, 256);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (p->cursor) (void)hipFree(p->cursor);
        delete p;
        set_error(std::string("plan create: ") + hipGetErrorString(e));
        return PYSIT_AMD_EDEVICE;
    }
    *out = p;
    return PYSIT_AMD_OK;
}

int pysit_amd_plan_stats(const pysit_amd_plan *plan, long long *stats)
{
    clear_error();
    PA_CHECK(plan && stats, "null argument");
    stats[0] = plan->n_launch;
    stats[1] = plan->n_capture;
    stats[2] = plan->n_direct;
    stats[3] = (long long)plan->graphs.size();
    return PYSIT_AMD_OK;
}

int pysit_amd_plan_destroy(pysit_amd_plan *plan)
{
    clear_error();
    if (!plan) return PYSIT_AMD_OK;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(plan->device);
    const hipError_t e = hipDeviceSynchronize();
    for (auto &g : plan->graphs) free_entry(g);
    if (plan->cap) (void)hipStreamDestroy(plan->cap);
    if (plan->cursor) (void)hipFree(plan->cursor);
    (void)hipSetDevice(cur);
    delete plan;
    if (e != hipSuccess) {
        set_error(std::string("plan destroy: ") + hipGetErrorString(e));
        return PYSIT_AMD_EDEVICE;
    }
    return PYSIT_AMD_OK;
}

}  // extern "C"
