// Host-side helpers shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/alsvinn_hip.h"
#include "avh_stage_inst.hpp"

struct avh_ctx {
    int device;
    // device scratch of the per-call Volume entry points (wave-speed maxima,
    // constraint flag), allocated on first use; `lock` serialises those calls
    std::mutex lock;
    void* scratch = nullptr;
};

namespace avh {

inline std::string& last_error_slot() {
    static thread_local std::string s;
    return s;
}

inline avh_status fail(avh_status st, const std::string& msg) {
    last_error_slot() = msg;
    return st;
}

#define AVH_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::avh::fail(AVH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

inline int nvars_of(int eq) {
    switch (eq) {
    case AVH_EQ_BURGERS: return 1;
    case AVH_EQ_EULER1: return 3;
    case AVH_EQ_EULER2: return 4;
    case AVH_EQ_EULER3: return 5;
    }
    return 0;
}

inline int ghosts_of(int recon) {
    switch (recon) {
    case AVH_RECON_NONE: return 1;
    case AVH_RECON_WENO2: return 2;
    case AVH_RECON_ENO2: return 2;
    case AVH_RECON_ENO3: return 3;
    }
    return 0;
}

inline Launchers launchers_for(int equation, int recon, int flux, int dim) {
    switch (equation) {
    case AVH_EQ_BURGERS: return get_launchers_burgers(recon, flux, dim);
    case AVH_EQ_EULER1: return get_launchers_euler1(recon, flux, dim);
    case AVH_EQ_EULER2: return get_launchers_euler2(recon, flux, dim);
    case AVH_EQ_EULER3: return get_launchers_euler3(recon, flux, dim);
    }
    return Launchers();
}

// Per-launch timing: an event after every launch of every step, all
// recorded before one synchronisation; nk launches and nk + 1 marks per step.
struct EventMarks {
    std::vector<hipEvent_t> ev;
    size_t count = 0;
    EventMarks() = default;
    EventMarks(const EventMarks&) = delete;
    EventMarks& operator=(const EventMarks&) = delete;
    ~EventMarks() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    avh_status create(size_t n) {
        ev.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            hipEvent_t e;
            AVH_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        return AVH_OK;
    }
    avh_status mark(hipStream_t st) {
        if (count >= ev.size()) return fail(AVH_EINVAL, "kernel timer overflow");
        AVH_HIP(hipEventRecord(ev[count++], st));
        return AVH_OK;
    }
    // ms[i] = mean duration of launch i over the steps
    avh_status average(int64_t steps, int nk, double* ms) {
        if (count != (size_t)steps * (nk + 1)) return fail(AVH_EINVAL, "kernel timer: unexpected launch count");
        AVH_HIP(hipEventSynchronize(ev[count - 1]));
        for (int i = 0; i < nk; ++i) ms[i] = 0;
        for (int64_t k = 0; k < steps; ++k)
            for (int i = 0; i < nk; ++i) {
                float f = 0;
                AVH_HIP(hipEventElapsedTime(&f, ev[k * (nk + 1) + i], ev[k * (nk + 1) + i + 1]));
                ms[i] += f;
            }
        for (int i = 0; i < nk; ++i) ms[i] /= (double)steps;
        return AVH_OK;
    }
};

inline avh_status mark(EventMarks* t, hipStream_t st) { return t ? t->mark(st) : AVH_OK; }

// A chunk of time steps captured once into a hipGraph and replayed. The step
// launches take fixed buffer pointers and read everything that changes (dt,
// time, save counters, the active flag) from the device control blocks, so a
// captured chunk stays valid for the solver's lifetime. No memset nodes are
// recorded: the end-of-step kernel clears what must start at zero.
struct StepGraph {
    int steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        steps = 0;
    }
    ~StepGraph() { reset(); }
};

// Enqueues n steps: whole chunks of graph_steps as graph launches (captured
// on first use), the remainder as ordinary launches. graph_steps < 2 turns
// graphs off.
// Captures and instantiates the chunk graph if it is not there yet; nothing
// is launched.
template <class Step>
avh_status ensure_step_graph(StepGraph& G, int graph_steps, hipStream_t st, Step&& step) {
    if (graph_steps < 2 || (G.steps == graph_steps && G.exec)) return AVH_OK;
    G.reset();
    AVH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    avh_status r = AVH_OK;
    for (int i = 0; i < graph_steps && r == AVH_OK; ++i) r = step();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (r != AVH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(AVH_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(AVH_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    }
    G.graph = g;
    G.exec = x;
    G.steps = graph_steps;
    return AVH_OK;
}

template <class Step>
avh_status enqueue_steps_graphed(StepGraph& G, int graph_steps, hipStream_t st, int64_t n, Step&& step) {
    if (graph_steps >= 2 && n >= graph_steps) {
        avh_status rg = ensure_step_graph(G, graph_steps, st, step);
        if (rg != AVH_OK) return rg;
        while (n >= graph_steps) {
            AVH_HIP(hipGraphLaunch(G.exec, st));
            n -= graph_steps;
        }
    }
    for (; n > 0; --n) {
        avh_status r = step();
        if (r != AVH_OK) return r;
    }
    return AVH_OK;
}

// Autotuner cache: the tile shape chosen for a (equation, reconstruction,
// flux, grid) of single solvers (kind 0) or batches (kind 1, grid = the
// first group, n[2] = total samples), shared by every solver of the process.
typedef std::tuple<int, int, int, int, long long, long long, long long> TuneKey;

inline std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
}

inline std::map<TuneKey, int>& tune_cache() {
    static std::map<TuneKey, int> c;
    return c;
}

inline bool tune_lookup(const TuneKey& k, int* shape) {
    std::lock_guard<std::mutex> g(tune_mutex());
    auto it = tune_cache().find(k);
    if (it == tune_cache().end()) return false;
    *shape = it->second;
    return true;
}

inline void tune_store(const TuneKey& k, int shape) {
    std::lock_guard<std::mutex> g(tune_mutex());
    tune_cache()[k] = shape;
}

// Mean milliseconds of `reps` back-to-back calls of launch(), after one
// untimed call, with HIP events on the launch stream.
template <class Launch>
avh_status time_launches(hipStream_t st, int reps, Launch&& launch, double* ms) {
    hipEvent_t e0, e1;
    AVH_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(AVH_EHIP, "hipEventCreate failed");
    }
    hipError_t e = launch();
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(AVH_EHIP, std::string("autotune timing: ") + hipGetErrorString(e));
    *ms = f / reps;
    return AVH_OK;
}

}  // namespace avh
