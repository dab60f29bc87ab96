// batch.hip -- batched ensemble of independent runs (C-ABI of
// include/nemo_gpu.h, nemo_gpu_batch_*), gfx950.
//
// Every run keeps its own mesh, state and CFL time step.  Runs of equal
// polynomial degree are concatenated into one group: each run's elements are
// padded to a multiple of the workgroup tile (BE elements) with ghost
// elements (own neighbour, constant state, element size 1e30, so they stay
// constant and never set a time step), so that a block of BE elements
// belongs to one run; a per-block run table gives the block its run's dt.
// Per RK stage one persistent launch per group walks all blocks of all its
// runs; the CFL partials of each block are reduced per run.  A pair of time
// steps (ten stage flips, so the ping-pong buffers return to their start) is
// captured as a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <vector>

#include "host_ops.hpp"
#include "dg_batch.hpp"

namespace {

using namespace nemo_host;

struct Group {
    int M = 0, BE = 0;
    std::vector<int> runs;
    long long nelem = 0, nblocks = 0, part_off = 0;
    double* u[2] = {nullptr, nullptr};
    double* r0 = nullptr;
    double* r1 = nullptr;
    double* delta = nullptr;
    int32_t* nbr = nullptr;
    int32_t* block_run = nullptr;
    int cur = 0;
};

struct Batch {
    nemo_gpu_config cfg{};
    int nruns = 0;
    std::vector<int> run_M, run_group;
    std::vector<long long> run_nelem, run_off;  // element offset inside the group (padded)
    Group g[3];                                 // M = 3, 4, 5
    HostOps ops[3];
    nemo::RKCoeffs rk{};
    double* dt_run = nullptr;   // [nruns]
    double* part = nullptr;     // CFL partials [all blocks][4]
    long long part_blocks = 0;
    int32_t* run_pstart = nullptr;  // [nruns] first partial of the run
    int32_t* run_pcount = nullptr;
    int32_t* run_M_dev = nullptr;
    int* status = nullptr;
    unsigned long long* step_ctr = nullptr;  // history row of the next dt
    double* hist = nullptr;
    long long hist_rows = 0;
    hipStream_t stream = nullptr;
    int ncu = 256;
    // launch geometry: one launch per stage for all degrees (one_launch) or
    // one per degree; grid_per_cu workgroups per CU walking the blocks
    // (0: one workgroup per block); the XCD-contiguous map; the order of the
    // degree groups in the one launch (0: M = 5, 4, 3, the costliest blocks
    // first; 1: M = 3, 4, 5)
    int grid_per_cu = 0, xcd = 1, one_launch = 1, order = 0;
    hipGraphExec_t graph = nullptr;
    int graph_cur[3] = {-1, -1, -1};
    int graph_grid = -1, graph_xcd = -1, graph_one = -1, graph_order = -1;
    // stage arguments of the one launch for the ten stages of two steps
    // (slot k: stage k % 5 + 1 from buffer parity k & 1), device memory
    void* margs = nullptr;
    int margs_order = -1;
    // timings of the last autotuning: (one_launch, order, grid_per_cu, xcd) -> ms per step
    std::vector<std::array<int, 4>> tune_geom;
    std::vector<double> tune_ms;
};

constexpr int group_of(int M) { return M - 3; }

template <int M>
constexpr int tile_be() {
    return Tile<M>::BE;
}

void free_batch(Batch* b) {
    for (auto& gr : b->g) {
        (void)hipFree(gr.u[0]);
        (void)hipFree(gr.u[1]);
        (void)hipFree(gr.r0);
        (void)hipFree(gr.r1);
        (void)hipFree(gr.delta);
        (void)hipFree(gr.nbr);
        (void)hipFree(gr.block_run);
    }
    (void)hipFree(b->dt_run);
    (void)hipFree(b->part);
    (void)hipFree(b->run_pstart);
    (void)hipFree(b->run_pcount);
    (void)hipFree(b->run_M_dev);
    (void)hipFree(b->status);
    (void)hipFree(b->step_ctr);
    (void)hipFree(b->hist);
    (void)hipFree(b->margs);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

template <int M, typename Real>
nemo::StageArgs<M, Real> batch_args(Batch* b, Group& gr, int stage, bool cfl) {
    nemo::StageArgs<M, Real> a{};
    a.u_in = gr.u[gr.cur];
    a.u_out = gr.u[gr.cur ^ 1];
    a.r0 = gr.r0;
    a.r1 = gr.r1;
    a.nbr = gr.nbr;
    a.delta = gr.delta;
    a.dt_run = b->dt_run;
    a.block_run = gr.block_run;
    a.status = b->status;
    a.cfl_part = cfl ? b->part + 4 * gr.part_off : nullptr;
    a.nelem = gr.nelem;
    a.stage = stage;
    a.timedisc = b->cfg.timedisc;
    a.indicator = b->cfg.indicator;
    a.repair = b->cfg.repair;
    a.alpha_max = b->cfg.alpha_max;
    a.alpha_min = b->cfg.alpha_min;
    a.ind_threshold = 0.5 * std::pow(10.0, -1.8 * std::pow(double(M), 0.25));
    a.ind_s = std::log((1.0 - 0.0001) / 0.0001);
    a.gas = make_gas<Real>(b->cfg.gamma);
    a.mhd = make_mhd<Real>(b->cfg.gamma, b->cfg.mu0);
    a.ph64 = make_phys64(b->cfg);
    a.op = make_table<M, Real>(b->ops[group_of(M)]);
    a.rk = b->rk;
    for (int i = 0; i < M; ++i) a.rk.w64[i] = b->ops[group_of(M)].w[i];
    a.alpha_refresh = 1;
    return a;
}

template <int M, typename Real>
void launch_group_stage(Batch* b, Group& gr, int stage, bool cfl) {
    constexpr int BE = Tile<M>::BE;
    const auto a = batch_args<M, Real>(b, gr, stage, cfl);
    if (b->grid_per_cu == 0) {
        // one workgroup per block: the plain stage kernel (its blocks are
        // dealt to the XCDs in contiguous stretches).  The persistent loop
        // keeps per-thread index arithmetic live across blocks, which costs
        // it about 120 VGPRs and most of its occupancy.
        hipLaunchKernelGGL((nemo::stage_kernel<nemo::EQ_EULER, M, BE, nemo::VOL_EC, nemo::SURF_HLLC, Real>),
                           dim3(unsigned(gr.nblocks)), dim3(BE * M * M), 0, b->stream, a);
        return;
    }
    long long grid = (long long)b->grid_per_cu * b->ncu;
    if (grid > gr.nblocks) grid = gr.nblocks;
    if (b->xcd) grid = (grid + 7) / 8 * 8;  // whole XCD rounds
    hipLaunchKernelGGL((nemo::stage_kernel_persistent<nemo::EQ_EULER, M, BE, nemo::VOL_EC, nemo::SURF_HLLC, Real>),
                       dim3(unsigned(grid)), dim3(BE * M * M), 0, b->stream, a, b->xcd);
}

template <int M>
void launch_group_cfl(Batch* b, Group& gr) {
    constexpr int BE = Tile<M>::BE;
    hipLaunchKernelGGL((nemo::cfl_kernel<nemo::EQ_EULER, M, BE>), dim3(unsigned(gr.nblocks)), dim3(BE * M * M), 0,
                       b->stream, gr.u[gr.cur], gr.delta, gr.nelem, make_phys64(b->cfg), b->part + 4 * gr.part_off);
}

// Per-run reduction of the block partials (one workgroup per run, fixed
// order: deterministic) and dt = 0.5*CFL/(2*N_NODES-1)*dtmin; the step
// counter selects the history row.
__global__ __launch_bounds__(256) void batch_finalize(const double* __restrict__ part, const int32_t* __restrict__ pstart,
                                                      const int32_t* __restrict__ pcount, const int32_t* __restrict__ runM,
                                                      double cfl, double* __restrict__ dt_run, double* __restrict__ hist,
                                                      long long hist_rows, int nruns,
                                                      const unsigned long long* __restrict__ ctr) {
    __shared__ double s[256];
    const int r = blockIdx.x, tid = threadIdx.x;
    double dt = 1.7976931348623157e308;
    for (int i = tid; i < pcount[r]; i += 256) {
        const double q = part[4 * ((long long)pstart[r] + i)];
        dt = q < dt ? q : dt;
    }
    s[tid] = dt;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) s[tid] = s[tid + off] < s[tid] ? s[tid + off] : s[tid];
        __syncthreads();
    }
    if (tid == 0) {
        const double d = 0.5 * cfl / double(2 * runM[r] - 1) * s[0];
        dt_run[r] = d;
        const unsigned long long row = *ctr;
        if (hist && row < (unsigned long long)hist_rows) hist[(long long)row * nruns + r] = d;
    }
}

__global__ void batch_count(unsigned long long* ctr) {
    if (threadIdx.x == 0) *ctr = *ctr + 1;
}

// Arguments of the one launch for the ten stages of two time steps: slot k
// is stage k % 5 + 1 reading buffer k & 1 (every stage flips all groups'
// buffers, and a step has five stages, so two steps return to the start).
template <typename Real>
int upload_multi_args(Batch* b) {
    std::vector<nemo::MultiStageArgs<Real>> h(10);
    int cur0[3];
    for (int i = 0; i < 3; ++i) cur0[i] = b->g[i].cur;
    for (int k = 0; k < 10; ++k) {
        for (int i = 0; i < 3; ++i) b->g[i].cur = k & 1;
        const int st = k % 5 + 1;
        auto& m = h[size_t(k)];
        if (b->g[0].nelem) m.a3 = batch_args<3, Real>(b, b->g[0], st, st == 5);
        if (b->g[1].nelem) m.a4 = batch_args<4, Real>(b, b->g[1], st, st == 5);
        if (b->g[2].nelem) m.a5 = batch_args<5, Real>(b, b->g[2], st, st == 5);
        for (int i = 0; i < 3; ++i) m.nblocks[i] = b->g[i].nblocks;
        const int ord[2][3] = {{2, 1, 0}, {0, 1, 2}};
        for (int i = 0; i < 3; ++i) m.order[i] = ord[b->order][i];
    }
    for (int i = 0; i < 3; ++i) b->g[i].cur = cur0[i];
    // earlier launches on the stream may still read the old arguments
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (!b->margs) HIP_TRY(hipMalloc(&b->margs, h.size() * sizeof(h[0])));
    HIP_TRY(hipMemcpy(b->margs, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
    b->margs_order = b->order;
    return NEMO_OK;
}

// Buffer parity shared by the non-empty groups (they flip together).
int batch_parity(const Batch* b) {
    for (const auto& gr : b->g)
        if (gr.nelem) return gr.cur;
    return 0;
}

// Device arguments of the one launch for the current order; called before
// any launch sequence (never during graph capture).
int prepare_launch(Batch* b) {
    if (!b->one_launch || b->margs_order == b->order) return NEMO_OK;
    return b->cfg.precision == 1 ? upload_multi_args<float>(b) : upload_multi_args<double>(b);
}

template <typename Real>
int launch_step_t(Batch* b) {
    if (b->one_launch) {
        // one launch per stage for the blocks of every degree
        if (b->margs_order != b->order) return fail(NEMO_ERR_NOT_READY, "batch launch arguments not prepared");
        const int k0 = batch_parity(b) ? 5 : 0;
        long long total = 0;
        for (const auto& gr : b->g) total += gr.nblocks;
        for (int st = 1; st <= 5; ++st) {
            const auto* pma = static_cast<const nemo::MultiStageArgs<Real>*>(b->margs) + (k0 + st - 1);
            if (b->grid_per_cu == 0) {
                hipLaunchKernelGGL((nemo::stage_kernel_multi<Real>), dim3(unsigned(total)), dim3(64), 0, b->stream,
                                   pma);
            } else {
                long long grid = (long long)b->grid_per_cu * b->ncu;
                if (grid > total) grid = total;
                if (b->xcd) grid = (grid + 7) / 8 * 8;  // whole XCD rounds
                hipLaunchKernelGGL((nemo::stage_kernel_multi_persistent<Real>), dim3(unsigned(grid)), dim3(64), 0,
                                   b->stream, pma, b->xcd);
            }
            for (auto& gr : b->g)
                if (gr.nelem) gr.cur ^= 1;
        }
    } else {
        for (int st = 1; st <= 5; ++st) {
            for (int gi = 0; gi < 3; ++gi) {
                Group& gr = b->g[gi];
                if (gr.nelem == 0) continue;
                switch (gr.M) {
                case 3: launch_group_stage<3, Real>(b, gr, st, st == 5); break;
                case 4: launch_group_stage<4, Real>(b, gr, st, st == 5); break;
                default: launch_group_stage<5, Real>(b, gr, st, st == 5); break;
                }
                gr.cur ^= 1;
            }
        }
    }
    // the dt of the next step goes to the next history row
    hipLaunchKernelGGL(batch_count, dim3(1), dim3(64), 0, b->stream, b->step_ctr);
    hipLaunchKernelGGL(batch_finalize, dim3(unsigned(b->nruns)), dim3(256), 0, b->stream, b->part, b->run_pstart,
                       b->run_pcount, b->run_M_dev, b->cfg.cfl, b->dt_run, b->hist, b->hist_rows, b->nruns,
                       b->step_ctr);
    HIP_TRY(hipGetLastError());
    return NEMO_OK;
}

int launch_step(Batch* b) {
    return b->cfg.precision == 1 ? launch_step_t<float>(b) : launch_step_t<double>(b);
}

// dt of the current states (first step of a run call).
int launch_cfl(Batch* b) {
    for (auto& gr : b->g) {
        if (gr.nelem == 0) continue;
        switch (gr.M) {
        case 3: launch_group_cfl<3>(b, gr); break;
        case 4: launch_group_cfl<4>(b, gr); break;
        default: launch_group_cfl<5>(b, gr); break;
        }
    }
    hipLaunchKernelGGL(batch_finalize, dim3(unsigned(b->nruns)), dim3(256), 0, b->stream, b->part, b->run_pstart,
                       b->run_pcount, b->run_M_dev, b->cfg.cfl, b->dt_run, b->hist, b->hist_rows, b->nruns,
                       b->step_ctr);
    HIP_TRY(hipGetLastError());
    return NEMO_OK;
}

bool graph_matches(const Batch* b) {
    if (!b->graph || b->graph_grid != b->grid_per_cu || b->graph_xcd != b->xcd || b->graph_one != b->one_launch ||
        b->graph_order != b->order)
        return false;
    for (int i = 0; i < 3; ++i)
        if (b->g[i].nelem && b->graph_cur[i] != b->g[i].cur) return false;
    return true;
}

// Capture two time steps (ten stage flips: the buffers return to their
// start) as one graph.
int build_graph(Batch* b) {
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    int cur0[3];
    for (int i = 0; i < 3; ++i) cur0[i] = b->g[i].cur;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(b);
    if (rc == NEMO_OK) rc = launch_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    for (int i = 0; i < 3; ++i) b->g[i].cur = cur0[i];  // capturing ran nothing
    if (rc != NEMO_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != NEMO_OK) return rc;
        return fail(NEMO_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(NEMO_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    for (int i = 0; i < 3; ++i) b->graph_cur[i] = b->g[i].cur;
    b->graph_grid = b->grid_per_cu;
    b->graph_xcd = b->xcd;
    b->graph_one = b->one_launch;
    b->graph_order = b->order;
    return NEMO_OK;
}

int run_steps(Batch* b, int nsteps, bool use_graph) {
    if (nsteps <= 0) return NEMO_OK;
    if (int rc = prepare_launch(b)) return rc;
    if (int rc = launch_cfl(b)) return rc;
    int it = 0;
    if (use_graph && nsteps >= 2) {
        if (!graph_matches(b))
            if (int rc = build_graph(b)) return rc;
        for (; it + 2 <= nsteps; it += 2) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    }
    for (; it < nsteps; ++it)
        if (int rc = launch_step(b)) return rc;
    return NEMO_OK;
}

int status_check(Batch* b) {
    int code = 0;
    HIP_TRY(hipMemcpyAsync(&code, b->status, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (code) {
        HIP_TRY(hipMemsetAsync(b->status, 0, sizeof(int), b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        return fail(NEMO_ERR_STATE, "positivity repair failed in a batched run");
    }
    return NEMO_OK;
}

}  // namespace

extern "C" {

int nemo_gpu_batch_create(const nemo_gpu_config* cfg, int32_t nruns, const int32_t* n_poly, const int64_t* nelem,
                          void** handle) {
    if (!cfg || !n_poly || !nelem || !handle || nruns <= 0) return fail(NEMO_ERR_ARG, "invalid arguments");
    if (cfg->equations != NEMO_EQ_EULER || cfg->volume != NEMO_VOL_EC || cfg->surface != NEMO_SURF_HLLC)
        return fail(NEMO_ERR_ARG, "the batched ensemble is built for Euler with the EC volume flux and HLLC");
    if (cfg->indicator != NEMO_IND_NONE && cfg->indicator != NEMO_IND_PP)
        return fail(NEMO_ERR_ARG, "batched ensemble: indicator must be none or Persson-Peraire");
    if (cfg->timedisc < 0 || cfg->timedisc > 1 || cfg->precision < 0 || cfg->precision > 1 || !(cfg->gamma > 1.0))
        return fail(NEMO_ERR_ARG, "invalid configuration");
    for (int r = 0; r < nruns; ++r)
        if (n_poly[r] < 2 || n_poly[r] > 4 || nelem[r] <= 0) return fail(NEMO_ERR_ARG, "run with invalid n_poly or size");
    Batch* b = new Batch();
    b->cfg = *cfg;
    b->nruns = nruns;
    b->rk = make_rk();
    for (int M = 3; M <= 5; ++M) {
        build_host_ops(M - 1, b->ops[group_of(M)]);
        b->g[group_of(M)].M = M;
    }
    b->g[0].BE = tile_be<3>();
    b->g[1].BE = tile_be<4>();
    b->g[2].BE = tile_be<5>();
    b->run_M.resize(nruns);
    b->run_group.resize(nruns);
    b->run_nelem.resize(nruns);
    b->run_off.resize(nruns);
    std::vector<int32_t> pstart(nruns), pcount(nruns);
    for (int r = 0; r < nruns; ++r) {
        const int M = n_poly[r] + 1;
        Group& gr = b->g[group_of(M)];
        b->run_M[r] = M;
        b->run_group[r] = group_of(M);
        b->run_nelem[r] = nelem[r];
        b->run_off[r] = gr.nelem;
        gr.runs.push_back(r);
        gr.nelem += (nelem[r] + gr.BE - 1) / gr.BE * gr.BE;
    }
    long long poff = 0;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        b->ncu = prop.multiProcessorCount;
    auto bail = [&](const std::string& msg) {
        free_batch(b);
        delete b;
        return fail(NEMO_ERR_HIP, msg);
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail("hipStreamCreate failed");
    for (auto& gr : b->g) {
        gr.nblocks = gr.nelem / (gr.BE > 0 ? gr.BE : 1);
        gr.part_off = poff;
        poff += gr.nblocks;
        if (gr.nelem == 0) continue;
        const size_t sb = size_t(5) * size_t(gr.nelem) * size_t(gr.M * gr.M * gr.M) * sizeof(double);
        if (hipMalloc(&gr.u[0], sb) != hipSuccess || hipMalloc(&gr.u[1], sb) != hipSuccess ||
            hipMalloc(&gr.r0, sb) != hipSuccess || hipMalloc(&gr.r1, sb) != hipSuccess ||
            hipMalloc(&gr.delta, size_t(gr.nelem) * 3 * sizeof(double)) != hipSuccess ||
            hipMalloc(&gr.nbr, size_t(gr.nelem) * 6 * sizeof(int32_t)) != hipSuccess ||
            hipMalloc(&gr.block_run, size_t(gr.nblocks) * sizeof(int32_t)) != hipSuccess)
            return bail("hipMalloc of the batch state failed");
        // ghost elements everywhere until the runs upload their meshes and states
        const int M3 = gr.M * gr.M * gr.M;
        std::vector<int32_t> nbr(size_t(gr.nelem) * 6), brun(size_t(gr.nblocks));
        std::vector<double> delta(size_t(gr.nelem) * 3, 1.0e30);
        for (long long e = 0; e < gr.nelem; ++e)
            for (int f = 0; f < 6; ++f) nbr[size_t(e) * 6 + f] = int32_t(e);
        for (int r : gr.runs) {
            const long long b0 = b->run_off[r] / gr.BE, nb = (b->run_nelem[r] + gr.BE - 1) / gr.BE;
            for (long long k = 0; k < nb; ++k) brun[size_t(b0 + k)] = r;
            pstart[r] = int32_t(gr.part_off + b0);
            pcount[r] = int32_t(nb);
        }
        std::vector<double> ghost(size_t(gr.nelem) * M3);
        bool ok = hipMemcpy(gr.nbr, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(gr.delta, delta.data(), delta.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(gr.block_run, brun.data(), brun.size() * sizeof(int32_t), hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemset(gr.r0, 0, sb) == hipSuccess && hipMemset(gr.r1, 0, sb) == hipSuccess;
        const double vals[5] = {1.0, 0.0, 0.0, 0.0, 1.0 / (cfg->gamma - 1.0)};
        for (int v = 0; v < 5 && ok; ++v) {
            std::fill(ghost.begin(), ghost.end(), vals[v]);
            ok = hipMemcpy(gr.u[0] + size_t(v) * gr.nelem * M3, ghost.data(), ghost.size() * sizeof(double),
                           hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(gr.u[1] + size_t(v) * gr.nelem * M3, ghost.data(), ghost.size() * sizeof(double),
                           hipMemcpyHostToDevice) == hipSuccess;
        }
        if (!ok) return bail("batch initialisation copies failed");
    }
    b->part_blocks = poff;
    std::vector<int32_t> rm(b->run_M.begin(), b->run_M.end());
    if (hipMalloc(&b->dt_run, size_t(nruns) * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->part, size_t(poff > 0 ? poff : 1) * 4 * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->run_pstart, size_t(nruns) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&b->run_pcount, size_t(nruns) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&b->run_M_dev, size_t(nruns) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&b->status, sizeof(int)) != hipSuccess || hipMalloc(&b->step_ctr, sizeof(unsigned long long)) != hipSuccess ||
        hipMemcpy(b->run_pstart, pstart.data(), size_t(nruns) * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->run_pcount, pcount.data(), size_t(nruns) * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->run_M_dev, rm.data(), size_t(nruns) * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->status, 0, sizeof(int)) != hipSuccess || hipMemset(b->step_ctr, 0, sizeof(unsigned long long)) != hipSuccess)
        return bail("hipMalloc of the batch tables failed");
    *handle = b;
    return NEMO_OK;
}

int nemo_gpu_batch_destroy(void* handle) {
    Batch* b = static_cast<Batch*>(handle);
    if (!b) return NEMO_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    delete b;
    return NEMO_OK;
}

int nemo_gpu_batch_upload(void* handle, int32_t run, const int32_t* nbr, const double* delta, const double* u) {
    Batch* b = static_cast<Batch*>(handle);
    if (!b || run < 0 || run >= b->nruns || !nbr || !delta || !u) return fail(NEMO_ERR_ARG, "invalid arguments");
    Group& gr = b->g[b->run_group[run]];
    const long long ne = b->run_nelem[run], off = b->run_off[run];
    const int M3 = gr.M * gr.M * gr.M;
    std::vector<int32_t> n(size_t(ne) * 6);
    for (long long i = 0; i < 6 * ne; ++i) {
        if (nbr[i] < 0 || nbr[i] >= ne) return fail(NEMO_ERR_ARG, "neighbour index out of range");
        n[size_t(i)] = int32_t(nbr[i] + off);
    }
    for (long long i = 0; i < 3 * ne; ++i)
        if (!(delta[i] > 0.0)) return fail(NEMO_ERR_ARG, "element size must be positive");
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(gr.nbr + off * 6, n.data(), n.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(gr.delta + off * 3, delta, size_t(ne) * 3 * sizeof(double), hipMemcpyHostToDevice));
    std::vector<double> col(size_t(ne) * M3);
    for (int v = 0; v < 5; ++v) {
        for (long long e = 0; e < ne; ++e)
            std::memcpy(&col[size_t(e) * M3], &u[(size_t(e) * 5 + v) * M3], M3 * sizeof(double));
        HIP_TRY(hipMemcpy(gr.u[gr.cur] + (size_t(v) * gr.nelem + off) * M3, col.data(), col.size() * sizeof(double),
                          hipMemcpyHostToDevice));
    }
    return NEMO_OK;
}

int nemo_gpu_batch_download(void* handle, int32_t run, double* u) {
    Batch* b = static_cast<Batch*>(handle);
    if (!b || run < 0 || run >= b->nruns || !u) return fail(NEMO_ERR_ARG, "invalid arguments");
    Group& gr = b->g[b->run_group[run]];
    const long long ne = b->run_nelem[run], off = b->run_off[run];
    const int M3 = gr.M * gr.M * gr.M;
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<double> col(size_t(ne) * M3);
    for (int v = 0; v < 5; ++v) {
        HIP_TRY(hipMemcpy(col.data(), gr.u[gr.cur] + (size_t(v) * gr.nelem + off) * M3, col.size() * sizeof(double),
                          hipMemcpyDeviceToHost));
        for (long long e = 0; e < ne; ++e)
            std::memcpy(&u[(size_t(e) * 5 + v) * M3], &col[size_t(e) * M3], M3 * sizeof(double));
    }
    return NEMO_OK;
}

int nemo_gpu_batch_set_launch(void* handle, int32_t grid_per_cu, int32_t xcd_map, int32_t one_launch, int32_t order) {
    Batch* b = static_cast<Batch*>(handle);
    if (!b || grid_per_cu < 0 || grid_per_cu > 64) return fail(NEMO_ERR_ARG, "grid_per_cu must be in [0, 64]");
    if (order < 0 || order > 1) return fail(NEMO_ERR_ARG, "order must be 0 or 1");
    b->grid_per_cu = grid_per_cu;
    b->xcd = xcd_map ? 1 : 0;
    b->one_launch = one_launch ? 1 : 0;
    b->order = order;
    return NEMO_OK;
}

int nemo_gpu_batch_run(void* handle, int32_t nsteps, int32_t use_graph, double* dt_hist) {
    Batch* b = static_cast<Batch*>(handle);
    if (!b || nsteps < 0) return fail(NEMO_ERR_ARG, "invalid arguments");
    if (dt_hist && nsteps > 0) {
        if (b->hist_rows < nsteps) {
            (void)hipFree(b->hist);
            b->hist = nullptr;
            b->hist_rows = 0;
            if (b->graph) {  // the graph holds the old history pointer
                (void)hipGraphExecDestroy(b->graph);
                b->graph = nullptr;
            }
            HIP_TRY(hipMalloc(&b->hist, size_t(nsteps) * b->nruns * sizeof(double)));
            b->hist_rows = nsteps;
        }
        HIP_TRY(hipMemsetAsync(b->step_ctr, 0, sizeof(unsigned long long), b->stream));
    } else {
        // no history: park the counter far past the rows
        const unsigned long long big = 1ull << 62;
        HIP_TRY(hipMemcpyAsync(b->step_ctr, &big, sizeof(big), hipMemcpyHostToDevice, b->stream));
    }
    if (int rc = run_steps(b, nsteps, use_graph != 0)) return rc;
    if (dt_hist && nsteps > 0)
        HIP_TRY(hipMemcpyAsync(dt_hist, b->hist, size_t(nsteps) * b->nruns * sizeof(double), hipMemcpyDeviceToHost,
                               b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return status_check(b);
}

int nemo_gpu_batch_autotune(void* handle, int32_t steps, int32_t* grid_per_cu, int32_t* xcd_map, int32_t* one_launch,
                            int32_t* order, double* best_ms_per_step) {
    Batch* b = static_cast<Batch*>(handle);
    if (!b || steps <= 0) return fail(NEMO_ERR_ARG, "invalid arguments");
    // snapshot the states (the tuning steps must not advance the runs)
    std::vector<double*> snap(3, nullptr);
    int cur0[3];
    for (int i = 0; i < 3; ++i) {
        Group& gr = b->g[i];
        cur0[i] = gr.cur;
        if (!gr.nelem) continue;
        const size_t sb = size_t(5) * gr.nelem * size_t(gr.M * gr.M * gr.M) * sizeof(double);
        HIP_TRY(hipMalloc(&snap[i], sb));
        HIP_TRY(hipMemcpyAsync(snap[i], gr.u[gr.cur], sb, hipMemcpyDeviceToDevice, b->stream));
    }
    const unsigned long long big = 1ull << 62;
    HIP_TRY(hipMemcpyAsync(b->step_ctr, &big, sizeof(big), hipMemcpyHostToDevice, b->stream));
    // candidates: one launch per stage (both group orders) or one per
    // degree, each with one workgroup per block or a persistent grid of 8
    // workgroups per CU with and without the XCD map (smaller persistent
    // grids leave the CUs underoccupied: 2-4x slower when measured)
    const int grids[] = {0, 8};
    struct Cand { int one, order, grid, x; };
    std::vector<Cand> cands;
    for (int one = 1; one >= 0; --one)
        for (int ord = 0; ord < (one ? 2 : 1); ++ord)
            for (int gpc : grids)
                for (int x = 0; x < 2; ++x)
                    if (!(gpc == 0 && x == 1)) cands.push_back({one, ord, gpc, x});
    double best = 1e300;
    int bg = b->grid_per_cu, bx = b->xcd, bone = b->one_launch, bord = b->order;
    b->tune_geom.clear();
    b->tune_ms.clear();
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = NEMO_OK;
    for (const Cand& c : cands) {
        if (rc != NEMO_OK) break;
        {
            b->grid_per_cu = c.grid;
            b->xcd = c.x;
            b->one_launch = c.one;
            b->order = c.order;
            const int gpc = c.grid, x = c.x;
            rc = prepare_launch(b);
            if (rc == NEMO_OK) rc = launch_cfl(b);
            if (rc == NEMO_OK) rc = launch_step(b);  // warm-up
            if (rc != NEMO_OK) break;
            (void)hipEventRecord(e0, b->stream);
            for (int s = 0; s < steps && rc == NEMO_OK; ++s) rc = launch_step(b);
            (void)hipEventRecord(e1, b->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double per = double(ms) / steps;
            b->tune_geom.push_back({c.one, c.order, c.grid, c.x});
            b->tune_ms.push_back(per);
            if (per < best) {
                best = per;
                bg = gpc;
                bx = x;
                bone = c.one;
                bord = c.order;
            }
            // back to the snapshot
            for (int i = 0; i < 3; ++i) {
                Group& gr = b->g[i];
                if (!gr.nelem) continue;
                gr.cur = cur0[i];
                const size_t sb = size_t(5) * gr.nelem * size_t(gr.M * gr.M * gr.M) * sizeof(double);
                (void)hipMemcpyAsync(gr.u[gr.cur], snap[i], sb, hipMemcpyDeviceToDevice, b->stream);
            }
        }
    }
    (void)hipStreamSynchronize(b->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto p : snap) (void)hipFree(p);
    HIP_TRY(hipMemsetAsync(b->status, 0, sizeof(int), b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (rc != NEMO_OK) return rc;
    b->grid_per_cu = bg;
    b->xcd = bx;
    b->one_launch = bone;
    b->order = bord;
    if (grid_per_cu) *grid_per_cu = bg;
    if (xcd_map) *xcd_map = bx;
    if (one_launch) *one_launch = bone;
    if (order) *order = bord;
    if (best_ms_per_step) *best_ms_per_step = best;
    return NEMO_OK;
}

int nemo_gpu_batch_tune_table(void* handle, int32_t max_rows, int32_t* rows, int32_t* geometry, double* ms) {
    Batch* b = static_cast<Batch*>(handle);
    if (!b || max_rows < 0 || !rows) return fail(NEMO_ERR_ARG, "invalid arguments");
    const int n = int(b->tune_ms.size());
    *rows = n;
    for (int i = 0; i < n && i < max_rows; ++i) {
        if (geometry)
            for (int k = 0; k < 4; ++k) geometry[4 * i + k] = b->tune_geom[size_t(i)][size_t(k)];
        if (ms) ms[i] = b->tune_ms[size_t(i)];
    }
    return NEMO_OK;
}

}  // extern "C"
