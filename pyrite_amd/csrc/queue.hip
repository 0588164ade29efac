// queue.hip -- many beam/target configurations in one persistent work-queue
// transport kernel, optionally captured in a hipGraph.
//
// The reference runs one case at a time (montecarlo/runner/case_driver.py:
// 142-152, its device core per case through transport/batching.py:477-628); a
// campaign of uneven cases is a host loop of launches that each drain at the
// speed of their longest history.  Here every history of every configuration
// is one item of a global queue: item i belongs to configuration c with
// off[c] <= i < off[c+1] and is that configuration's electron e = i - off[c],
// with the reference's stream key splitmix64(seed_c + PHI (e+1)).  Lanes claim
// `per_claim` consecutive items with one aggregated atomic per wave (ballot +
// popcount), so configurations of 1e4 and 1e7 histories, thin and thick,
// share the grid and the tail of one case overlaps the body of the next.
// Histories run the analog step (transport_step.hpp) without storing rows;
// each history's outcome is tallied per configuration (FP64 atomics).
#include "transport_step.hpp"

namespace hr {
namespace {

constexpr int kQueueMaxCfg = 1024;
constexpr int kTally = HR_QUEUE_TALLY;

__device__ __forceinline__ int find_config(const int64_t *__restrict__ off, int n_cfg, int64_t i) {
    int lo = 0, hi = n_cfg - 1;  // last c with off[c] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Energies enter the tally quantised to 2^-12 keV (0.24 eV): FP64 sums of
// integer multiples of the quantum are exact up to 2^53 quanta (2.2e12 keV per
// configuration), so, like the integer-valued counts, they do not depend on
// the order in which the atomics land (the claim order): the tallies are
// bitwise reproducible.
constexpr double kTallyQuantum = 1.0 / 4096.0;
__device__ __forceinline__ double quantised(double keV) { return rint(keV * 4096.0) * kTallyQuantum; }

__device__ __forceinline__ void tally_history(double *__restrict__ t, const Lane &s, int code, double E0) {
    atomicAdd(&t[code], 1.0);
    atomicAdd(&t[7], (double)s.nseg);
    atomicAdd(&t[8], (double)s.nseg * (double)s.nseg);
    if (code == EXIT_TRANS) atomicAdd(&t[9], quantised(s.E));
    if (code == EXIT_BACK) atomicAdd(&t[10], quantised(s.E));
    const bool escaped = code == EXIT_BACK || code == EXIT_TRANS || code == EXIT_SIDE;
    atomicAdd(&t[11], quantised(E0) - (escaped ? quantised(s.E) : 0.0));
}

template <bool kMixed>
__global__ void __launch_bounds__(256) queue_kernel(const hr_queue_config *__restrict__ cfg, int32_t n_cfg,
                                                    const int64_t *__restrict__ off_g, int64_t total,
                                                    double *__restrict__ tally, unsigned long long *next,
                                                    int32_t per_claim) {
    __shared__ int64_t off[kQueueMaxCfg + 1];
    for (int i = threadIdx.x; i <= n_cfg; i += blockDim.x) off[i] = off_g[i];
    __syncthreads();
    Lane s;
    bool live = false, drained = false;
    int c = 0;
    int64_t id = 0, id_end = 0;
    while (true) {
        const bool want = !live && !drained && id >= id_end;
        const uint64_t mask = __ballot(want);
        if (mask != 0) {
            const int lane = __lane_id();
            const int leader = __ffsll((unsigned long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(next, (unsigned long long)__popcll(mask) * (unsigned long long)per_claim);
            base = __shfl(base, leader);
            if (want) {
                const int rank = __popcll(mask & ((1ULL << lane) - 1ULL));
                id = (int64_t)base + (int64_t)rank * per_claim;
                id_end = id + per_claim < total ? id + per_claim : total;
                if (id >= total) drained = true;
            }
        }
        if (!live && !drained && id < id_end) {
            c = find_config(off, n_cfg, id);
            const hr_queue_config &q = cfg[c];
            const int64_t e = id - off[c];
            id += 1;
            s.e = e;
            s.id = e;
            s.px = 0.0;
            s.py = 0.0;
            s.pz = 0.0;
            s.dx = q.beam_dir[0];
            s.dy = q.beam_dir[1];
            s.dz = q.beam_dir[2];
            s.E = q.E0_keV;
            s.clock = 0.0;
            s.E_cut = q.E_cut_keV;
            s.key = splitmix64(q.seed + kGolden * ((uint64_t)e + 1ULL));
            s.tau = -1.0;
            s.stragg = 0.0;
            s.draws = 0;
            s.nseg = 0;
            s.flight = 0;
            s.substep = 0;
            s.steps = 0;
            live = q.t.max_steps > 0;
            if (!live) tally_history(tally + (int64_t)c * kTally, s, EXIT_STEP_LIMITED, q.E0_keV);
        }
        if (__ballot(live || !drained) == 0) break;
        if (live) {
            const int code = step_lane<false, kMixed>(cfg[c].t, s);
            if (code >= 0) {
                tally_history(tally + (int64_t)c * kTally, s, code, cfg[c].E0_keV);
                live = false;
            }
        }
    }
}

__global__ void queue_reset_kernel(unsigned long long *next, double *__restrict__ tally, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *next = 0ULL;
    if (i < n) tally[i] = 0.0;
}

hr_status queue_check(const hr_queue_launch *L) {
    HR_REQUIRE(L != nullptr, "queue: null launch");
    HR_REQUIRE(L->n_cfg >= 1 && L->n_cfg <= kQueueMaxCfg, "queue: n_cfg %d outside [1, %d]", L->n_cfg,
               kQueueMaxCfg);
    HR_REQUIRE(L->configs && L->offsets && L->tally && L->counter, "queue: null pointer");
    HR_REQUIRE(L->total >= 0, "queue: negative total");
    HR_REQUIRE(L->per_claim >= 1 && L->per_claim <= 1024, "queue: per_claim %d outside [1, 1024]", L->per_claim);
    HR_REQUIRE(L->block_size == 0 || (L->block_size % 64 == 0 && L->block_size >= 64 && L->block_size <= 256),
               "queue: block_size must be 64, 128, 192 or 256 (or 0)");
    HR_REQUIRE(L->precision == 0 || L->precision == 1, "queue: precision must be 0 (fp64) or 1 (mixed)");
    return HR_OK;
}

// reset + queue kernel on `s` with the given counter
hr_status queue_enqueue(const hr_queue_launch *L, unsigned long long *counter, int cu, hipStream_t s) {
    const int block = L->block_size == 0 ? 256 : L->block_size;
    const int64_t nt = (int64_t)L->n_cfg * kTally;
    hipLaunchKernelGGL(queue_reset_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, counter, L->tally,
                       L->accumulate ? 0 : nt);
    HR_CHECK_HIP(hipGetLastError());
    if (L->total == 0) return HR_OK;
    int occ = 0;
    HR_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, queue_kernel<false>, block, 0));
    occ = occ > 0 ? occ : 1;
    const int64_t need = (L->total / L->per_claim + block) / block + 1;
    const int64_t resident = (int64_t)cu * occ;
    const int64_t grid = need < resident ? need : resident;
    if (L->precision == 1)
        hipLaunchKernelGGL(queue_kernel<true>, dim3((unsigned)grid), dim3(block), 0, s, L->configs, L->n_cfg,
                           L->offsets, L->total, L->tally, counter, L->per_claim);
    else
        hipLaunchKernelGGL(queue_kernel<false>, dim3((unsigned)grid), dim3(block), 0, s, L->configs, L->n_cfg,
                           L->offsets, L->total, L->tally, counter, L->per_claim);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

struct QueueGraph {
    hipGraphExec_t exec = nullptr;
};

}  // namespace
}  // namespace hr

extern "C" {

hr_status hr_queue_run(const hr_queue_launch *launch, void *stream) {
    HR_TRACE("hr_queue_run");
    using namespace hr;
    hr_status st = queue_check(launch);
    if (st != HR_OK) return st;
    int cu = 0;
    st = device_cu(cu);
    if (st != HR_OK) return st;
    return queue_enqueue(launch, launch->counter, cu, (hipStream_t)stream);
}

hr_status hr_queue_graph_create(const hr_queue_launch *launch, void **graph) {
    using namespace hr;
    HR_REQUIRE(graph != nullptr, "hr_queue_graph_create: null handle");
    *graph = nullptr;
    hr_status st = queue_check(launch);
    if (st != HR_OK) return st;
    int cu = 0;
    st = device_cu(cu);
    if (st != HR_OK) return st;
    QueueGraph *g = new QueueGraph();
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete g;
        HR_CHECK_HIP(e);
    }
    hipGraph_t graph_t = nullptr;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        st = queue_enqueue(launch, launch->counter, cu, cs);
        const hipError_t e2 = hipStreamEndCapture(cs, &graph_t);
        if (st == HR_OK && e2 != hipSuccess) e = e2;
    }
    if (e == hipSuccess && st == HR_OK) e = hipGraphInstantiate(&g->exec, graph_t, nullptr, nullptr, 0);
    if (graph_t) (void)hipGraphDestroy(graph_t);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess || st != HR_OK) {
        delete g;
        if (st != HR_OK) return st;
        HR_CHECK_HIP(e);
    }
    *graph = g;
    return HR_OK;
}

hr_status hr_queue_graph_launch(void *graph, void *stream) {
    HR_TRACE("hr_queue_graph_launch");
    using namespace hr;
    HR_REQUIRE(graph != nullptr, "hr_queue_graph_launch: null graph");
    QueueGraph *g = (QueueGraph *)graph;
    HR_CHECK_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return HR_OK;
}

hr_status hr_queue_graph_destroy(void *graph) {
    using namespace hr;
    if (graph == nullptr) return HR_OK;
    QueueGraph *g = (QueueGraph *)graph;
    if (g->exec) HR_CHECK_HIP(hipGraphExecDestroy(g->exec));
    delete g;
    return HR_OK;
}
}
