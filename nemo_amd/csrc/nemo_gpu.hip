// nemo_gpu.hip -- C-ABI implementation (include/nemo_gpu.h) for gfx950.
//
// Host side of the MI355X spatial operator: operator construction, device
// memory, the per-stage launch sequence and the error convention.  The device
// work lives in dg_stage.hpp; the stage launches of each polynomial degree
// are compiled apart (stage_launch.hpp, stage_m{3,4,5}.hip).  There is no CPU
// fallback: every compute entry point runs HIP kernels and reports HIP errors.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "../../include/nemo_gpu.h"
#include "host_ops.hpp"
#include "solver.hpp"
#include "dg_mortar.hpp"
#include "dg_analyze.hpp"

// Workgroup partials of nemo_gpu_analyze_noalloc: one row of
// NEMO_ANALYZE_NSUMS doubles per analysis workgroup of 256 nodes, enough for
// NEMO_ANALYZE_NOALLOC_MAX_NODES nodes per variable (8 MiB).  Every GPU the
// library runs on holds its own copy until the process ends; all calls on
// one GPU share it.  Host code never passes this name to a kernel: its device
// address comes from hipGetSymbolAddress (scratch_device_ptr below).
static __device__ __align__(256) double scratch[NEMO_ANALYZE_NOALLOC_MAX_NODES / nemo::AN_NT * nemo::AN_NSUMS];

namespace {

using namespace nemo_host;

// Device address of `scratch` on the current device, resolved once per
// device (with that device current) and kept for later calls.
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kMaxDevices] = {};

hipError_t scratch_device_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = scratch_dev[dev];
    return hipSuccess;
}


void free_mesh_buffers(Solver& s) {
    (void)hipFree(s.u[0]);
    (void)hipFree(s.u[1]);
    (void)hipFree(s.r0);
    (void)hipFree(s.r1);
    (void)hipFree(s.rhs);
    (void)hipFree(s.alpha);
    (void)hipFree(s.nbr);
    (void)hipFree(s.delta);
    (void)hipFree(s.part);
    (void)hipFree(s.alpha_step);
    s.u[0] = s.u[1] = s.r0 = s.r1 = s.rhs = s.alpha = s.delta = s.part = s.alpha_step = nullptr;
    s.nbr = nullptr;
    s.cap = 0;
    (void)hipFree(s.mortars);
    (void)hipFree(s.mortar_flux);
    (void)hipFree(s.bfaces);
    s.mortars = nullptr;
    s.mortar_flux = nullptr;
    s.bfaces = nullptr;
    s.nmortar = 0;
    s.nbound = 0;
    s.nslots = 0;
}


int launch_stage(Solver& s, int stage, double dt, const double* dt_dev, double* rhs_out, double* alpha_out,
                 double* cfl_part = nullptr, const LaunchOpts& o = LaunchOpts()) {
    if (s.cfg.indicator == NEMO_IND_MLP && !s.mlp_ready && !o.feat_out)
        return fail(NEMO_ERR_NOT_READY, "learned indicator: no network parameters (nemo_gpu_mlp_set_params or _train)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s.timing) {
        if (s.ev_used == s.ev.size()) {
            hipEvent_t a, b;
            HIP_TRY(hipEventCreate(&a));
            HIP_TRY(hipEventCreate(&b));
            s.ev.emplace_back(a, b);
        }
        e0 = s.ev[s.ev_used].first;
        e1 = s.ev[s.ev_used].second;
        ++s.ev_used;
        HIP_TRY(hipEventRecord(e0, s.stream));
    }
    int rc;
    switch (s.M) {
    case 3: rc = launch_stage_m<3>(s, stage, dt, dt_dev, rhs_out, alpha_out, cfl_part, o); break;
    case 4: rc = launch_stage_m<4>(s, stage, dt, dt_dev, rhs_out, alpha_out, cfl_part, o); break;
    case 5: rc = launch_stage_m<5>(s, stage, dt, dt_dev, rhs_out, alpha_out, cfl_part, o); break;
    default: return fail(NEMO_ERR_ARG, "n_poly must be 2, 3 or 4");
    }
    if (rc != NEMO_OK) return rc;
    HIP_TRY(hipGetLastError());
    if (s.timing) HIP_TRY(hipEventRecord(e1, s.stream));
    return NEMO_OK;
}

template <int M>
long long stage_blocks(const Solver& s) {
    return (s.nelem + Tile<M>::BE - 1) / Tile<M>::BE;
}

long long nblocks_for(const Solver& s) {
    switch (s.M) {
    case 3: return stage_blocks<3>(s);
    case 4: return stage_blocks<4>(s);
    default: return stage_blocks<5>(s);
    }
}

template <int M>
void launch_cfl_m(Solver& s) {
    constexpr int BE = Tile<M>::BE;
    const nemo::Phys64 P = make_phys64(s.cfg);
    if (s.cfg.equations == NEMO_EQ_GLM_MHD)
        hipLaunchKernelGGL((nemo::cfl_kernel<nemo::EQ_MHD, M, BE>), dim3(unsigned(stage_blocks<M>(s))),
                           dim3(BE * M * M), 0, s.stream, s.u[s.cur], s.delta, s.nelem, P, s.part);
    else
        hipLaunchKernelGGL((nemo::cfl_kernel<nemo::EQ_EULER, M, BE>), dim3(unsigned(stage_blocks<M>(s))),
                           dim3(BE * M * M), 0, s.stream, s.u[s.cur], s.delta, s.nelem, P, s.part);
}

int launch_finalize(Solver& s, long long hist_slot) {
    const long long npart = nblocks_for(s);
    long long G = (npart + nemo::CFL_NT - 1) / nemo::CFL_NT;
    G = G < 1 ? 1 : (G > nemo::CFL_GRID ? nemo::CFL_GRID : G);
    hipLaunchKernelGGL(nemo::cfl_finalize, dim3(unsigned(G)), dim3(nemo::CFL_NT), 0, s.stream, s.part, npart,
                       s.cfg.cfl, s.M, s.dt_dev, hist_slot >= 0 ? s.hist : nullptr, hist_slot, s.cfl_red,
                       s.cfl_ticket);
    HIP_TRY(hipGetLastError());
    return NEMO_OK;
}

// CFL time step of the current state into dt_dev (and hist[hist_slot]).
int launch_cfl(Solver& s, long long hist_slot) {
    switch (s.M) {
    case 3: launch_cfl_m<3>(s); break;
    case 4: launch_cfl_m<4>(s); break;
    case 5: launch_cfl_m<5>(s); break;
    default: return fail(NEMO_ERR_ARG, "n_poly must be 2, 3 or 4");
    }
    HIP_TRY(hipGetLastError());
    return launch_finalize(s, hist_slot);
}

int check_ready(Solver* s) {
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    if (!s->have_mesh) return fail(NEMO_ERR_NOT_READY, "mesh not uploaded");
    if (!s->have_state) return fail(NEMO_ERR_NOT_READY, "state not uploaded");
    return NEMO_OK;
}

// host layout u[e][v][n]  <->  device layout u[v][e][n]
void to_device_layout(const Solver& s, const double* host, std::vector<double>& dev) {
    dev.resize(s.state_doubles());
    const size_t M3 = s.M3, E = s.nelem;
    for (size_t e = 0; e < E; ++e)
        for (int v = 0; v < s.V; ++v)
            std::memcpy(&dev[(v * E + e) * M3], &host[(e * s.V + v) * M3], M3 * sizeof(double));
}

void to_host_layout(const Solver& s, const std::vector<double>& dev, double* host) {
    const size_t M3 = s.M3, E = s.nelem;
    for (size_t e = 0; e < E; ++e)
        for (int v = 0; v < s.V; ++v)
            std::memcpy(&host[(e * s.V + v) * M3], &dev[(v * E + e) * M3], M3 * sizeof(double));
}

int status_check(Solver* s) {
    int code = 0;
    HIP_TRY(hipMemcpyAsync(&code, s->status, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (code) {
        HIP_TRY(hipMemsetAsync(s->status, 0, sizeof(int), s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        return fail(NEMO_ERR_STATE, "positivity repair failed (kernel_repair returned false)");
    }
    return NEMO_OK;
}

// Device allocation released when it goes out of scope unless handed over.
struct DeviceBuffer {
    void* p = nullptr;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
};

// Mesh checks shared by the uploads: conforming neighbours in range, every
// non-conforming or boundary face pointing at the one face slot that
// belongs to it (mortar slots first, then the boundary faces').
int check_mesh(const Solver& s, int64_t nelem, const int32_t* nbr, const double* delta, int64_t nmortar,
               const int32_t* mortars, int64_t nbound = 0, const int32_t* bfaces = nullptr) {
    if (!nbr || !delta || nelem <= 0 || nelem > INT32_MAX) return fail(NEMO_ERR_ARG, "invalid mesh arguments");
    if (nmortar < 0 || (nmortar > 0 && !mortars) || nmortar > INT32_MAX / 5)
        return fail(NEMO_ERR_ARG, "invalid mortar arguments");
    if (nbound < 0 || (nbound > 0 && !bfaces) || nbound > INT32_MAX - 5 * nmortar)
        return fail(NEMO_ERR_ARG, "invalid boundary arguments");
    const int64_t nslots = 5 * nmortar + nbound;
    std::vector<int32_t> owner(size_t(nslots), -1);
    for (int64_t i = 0; i < 6 * nelem; ++i) {
        const int32_t n = nbr[i];
        if (n >= nelem) return fail(NEMO_ERR_ARG, "neighbour index out of range");
        if (n < 0) {
            const int64_t slot = -int64_t(n) - 1;
            if (slot >= nslots) return fail(NEMO_ERR_ARG, "face slot out of range");
            if (owner[size_t(slot)] != -1) return fail(NEMO_ERR_ARG, "face slot used by two faces");
            owner[size_t(slot)] = int32_t(i);
        }
    }
    for (int64_t b = 0; b < nbound; ++b) {
        const int32_t* t = bfaces + 3 * b;
        if (t[0] < 0 || t[0] >= nelem || t[1] < 0 || t[1] > 5) return fail(NEMO_ERR_ARG, "boundary face out of range");
        if (t[2] != nemo::BC_TRANSMISSIVE && t[2] != nemo::BC_SLIP_WALL)
            return fail(NEMO_ERR_ARG, "unknown boundary kind");
        if (t[2] == nemo::BC_SLIP_WALL && s.cfg.equations != NEMO_EQ_EULER)
            return fail(NEMO_ERR_ARG, "the slip wall is built for the Euler equations");
        if (owner[size_t(5 * nmortar + b)] != t[0] * 6 + t[1])
            return fail(NEMO_ERR_ARG, "boundary face does not match nbr");
    }
    for (int64_t m = 0; m < nmortar; ++m) {
        const int32_t* t = mortars + 6 * m;
        if (t[1] < 0 || t[1] > 5) return fail(NEMO_ERR_ARG, "mortar face out of range");
        for (int k = 0; k < 6; ++k)
            if (k != 1 && (t[k] < 0 || t[k] >= nelem)) return fail(NEMO_ERR_ARG, "mortar element out of range");
        if (owner[size_t(5 * m)] != t[0] * 6 + t[1]) return fail(NEMO_ERR_ARG, "mortar big face does not match nbr");
        for (int k = 0; k < 4; ++k)
            if (owner[size_t(5 * m + 1 + k)] != t[2 + k] * 6 + (t[1] ^ 1))
                return fail(NEMO_ERR_ARG, "mortar small face does not match nbr");
    }
    for (int64_t i = 0; i < 3 * nelem; ++i)
        if (!(delta[i] > 0.0)) return fail(NEMO_ERR_ARG, "element size must be positive");
    return NEMO_OK;
}

// Device copies of a checked mesh; state-sized buffers are reallocated when
// the element count changes (the state itself then needs an upload or a remap).
// The solver counts as having no mesh until every copy is in place, so a
// failure part way leaves no entry point running on half-built buffers.
int set_mesh(Solver& s, int64_t nelem, const int32_t* nbr, const double* delta, int64_t nmortar,
             const int32_t* mortars, int64_t nbound = 0, const int32_t* bfaces = nullptr) {
    HIP_TRY(hipStreamSynchronize(s.stream));
    s.have_mesh = false;
    if (nelem != s.nelem) {
        s.have_state = false;
        if (nelem > s.cap) {
            // mesh-sized buffers for this mesh and later ones up to 5/4 of it
            // (adaptation changes the element count a little at a time)
            free_mesh_buffers(s);
            const long long cap = nelem + nelem / 4;
            s.nelem = cap;  // sizes below are for the capacity
            const size_t bytes = s.state_doubles() * sizeof(double);
            HIP_TRY(hipMalloc(&s.u[0], bytes));
            HIP_TRY(hipMalloc(&s.u[1], bytes));
            HIP_TRY(hipMalloc(&s.r0, bytes));
            HIP_TRY(hipMalloc(&s.r1, bytes));
            HIP_TRY(hipMalloc(&s.rhs, bytes));
            HIP_TRY(hipMalloc(&s.alpha, size_t(cap) * sizeof(double)));
            HIP_TRY(hipMalloc(&s.nbr, size_t(cap) * 6 * sizeof(int32_t)));
            HIP_TRY(hipMalloc(&s.delta, size_t(cap) * 3 * sizeof(double)));
            HIP_TRY(hipMalloc(&s.part, size_t(nblocks_for(s)) * 4 * sizeof(double)));
            HIP_TRY(hipMalloc(&s.alpha_step, size_t(cap) * sizeof(double)));
            s.cap = cap;
        }
        s.nelem = nelem;
        const size_t bytes = s.state_doubles() * sizeof(double);
        HIP_TRY(hipMemsetAsync(s.alpha_step, 0, size_t(nelem) * sizeof(double), s.stream));
        HIP_TRY(hipMemsetAsync(s.r0, 0, bytes, s.stream));
        HIP_TRY(hipMemsetAsync(s.r1, 0, bytes, s.stream));
    }
    if (nmortar != s.nmortar) {
        (void)hipFree(s.mortars);
        s.mortars = nullptr;
        s.nmortar = nmortar;
        if (nmortar > 0) HIP_TRY(hipMalloc(&s.mortars, size_t(nmortar) * 6 * sizeof(int32_t)));
    }
    if (nbound != s.nbound) {
        (void)hipFree(s.bfaces);
        s.bfaces = nullptr;
        s.nbound = nbound;
        if (nbound > 0) HIP_TRY(hipMalloc(&s.bfaces, size_t(nbound) * 3 * sizeof(int32_t)));
    }
    if (5 * nmortar + nbound != s.nslots) {
        (void)hipFree(s.mortar_flux);
        s.mortar_flux = nullptr;
        s.nslots = 5 * nmortar + nbound;
        if (s.nslots > 0)
            HIP_TRY(hipMalloc(&s.mortar_flux, size_t(s.nslots) * size_t(s.V) * size_t(s.M * s.M) * sizeof(double)));
    }
    if (!s.counter) HIP_TRY(hipMalloc(&s.counter, sizeof(int)));
    HIP_TRY(hipMemcpyAsync(s.nbr, nbr, size_t(nelem) * 6 * sizeof(int32_t), hipMemcpyHostToDevice, s.stream));
    HIP_TRY(hipMemcpyAsync(s.delta, delta, size_t(nelem) * 3 * sizeof(double), hipMemcpyHostToDevice, s.stream));
    if (nmortar > 0)
        HIP_TRY(hipMemcpyAsync(s.mortars, mortars, size_t(nmortar) * 6 * sizeof(int32_t), hipMemcpyHostToDevice,
                               s.stream));
    if (nbound > 0)
        HIP_TRY(hipMemcpyAsync(s.bfaces, bfaces, size_t(nbound) * 3 * sizeof(int32_t), hipMemcpyHostToDevice,
                               s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    s.have_mesh = true;
    return NEMO_OK;
}

template <int M>
void launch_analyze(Solver& s, long long nblocks, bool mhd, double* part) {
    nemo::AnalyzeArgs<M> a{};
    a.u = s.u[s.cur];
    a.delta = s.delta;
    a.part = part;
    a.nelem = s.nelem;
    a.ph = make_phys64(s.cfg);
    for (int i = 0; i < M; ++i) a.w[i] = s.ops.w[i];
    if (mhd)
        hipLaunchKernelGGL((nemo::analyze_kernel<nemo::EQ_MHD, M>), dim3(unsigned(nblocks)), dim3(nemo::AN_NT), 0,
                           s.stream, a);
    else
        hipLaunchKernelGGL((nemo::analyze_kernel<nemo::EQ_EULER, M>), dim3(unsigned(nblocks)), dim3(nemo::AN_NT), 0,
                           s.stream, a);
}

template <int M>
void launch_remap(Solver& s, double* u_new, const int32_t* d_units, int64_t nunits, int64_t nelem_new) {
    nemo::RemapArgs<M> a{};
    a.u_old = s.u[s.cur];
    a.u_new = u_new;
    a.units = d_units;
    a.nold = s.nelem;
    a.nnew = nelem_new;
    a.ph64 = make_phys64(s.cfg);
    mortar_ops(s.ops, &a.op.P[0][0][0], &a.op.R[0][0][0]);
    for (int i = 0; i < M; ++i) a.w[i] = s.ops.w[i];
    a.fallbacks = s.counter;
    const dim3 grid{unsigned(nunits)}, block{unsigned(nemo::RemapShape<M>::NT)};
    if (s.cfg.equations == NEMO_EQ_GLM_MHD)
        hipLaunchKernelGGL((nemo::remap_kernel<nemo::EQ_MHD, M>), grid, block, 0, s.stream, a);
    else
        hipLaunchKernelGGL((nemo::remap_kernel<nemo::EQ_EULER, M>), grid, block, 0, s.stream, a);
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

std::string& nemo_host::last_error() {
    static thread_local std::string msg;
    return msg;
}

extern "C" {

int nemo_gpu_error_message(char* buf, int32_t len) {
    if (!buf || len <= 0) return NEMO_ERR_ARG;
    std::snprintf(buf, size_t(len), "%s", nemo_host::last_error().c_str());
    return NEMO_OK;
}

int nemo_gpu_operators(int32_t n_poly, double* x, double* w, double* D, double* Vinv) {
    if (n_poly < 1 || n_poly + 1 > kMaxM) return fail(NEMO_ERR_ARG, "n_poly out of range");
    HostOps op;
    build_host_ops(n_poly, op);
    const int M = op.M;
    if (x) std::memcpy(x, op.x, sizeof(double) * M);
    if (w) std::memcpy(w, op.w, sizeof(double) * M);
    if (D) std::memcpy(D, op.D, sizeof(double) * M * M);
    if (Vinv) std::memcpy(Vinv, op.Vinv, sizeof(double) * M * M);
    return NEMO_OK;
}

int nemo_gpu_create(const nemo_gpu_config* cfg, void** handle) {
    if (!cfg || !handle) return fail(NEMO_ERR_ARG, "null argument");
    if (cfg->n_poly < 2 || cfg->n_poly > 4) return fail(NEMO_ERR_ARG, "n_poly must be 2, 3 or 4");
    if (cfg->equations != NEMO_EQ_EULER && cfg->equations != NEMO_EQ_GLM_MHD)
        return fail(NEMO_ERR_ARG, "invalid equations");
    if (cfg->volume < 0 || cfg->volume > 2) return fail(NEMO_ERR_ARG, "invalid volume flux");
    if (cfg->equations == NEMO_EQ_EULER && (cfg->surface < 0 || cfg->surface > 1))
        return fail(NEMO_ERR_ARG, "invalid surface flux for Euler (Rusanov or HLLC)");
    if (cfg->equations == NEMO_EQ_GLM_MHD) {
        if (cfg->volume == NEMO_VOL_WEAK)
            return fail(NEMO_ERR_ARG, "GLM-MHD needs flux differencing (NEMO_VOL_CENTRAL or NEMO_VOL_EC)");
        if (cfg->surface != NEMO_SURF_RUSANOV && cfg->surface != NEMO_SURF_HLLD)
            return fail(NEMO_ERR_ARG, "invalid surface flux for GLM-MHD (Rusanov or HLLD)");
        if (!(cfg->mu0 > 0.0)) return fail(NEMO_ERR_ARG, "mu0 must be positive");
    }
    if (cfg->indicator < 0 || cfg->indicator > 2) return fail(NEMO_ERR_ARG, "invalid indicator");
    if (cfg->indicator == NEMO_IND_MLP &&
        (cfg->equations != NEMO_EQ_EULER || cfg->n_poly != 3 || cfg->volume != NEMO_VOL_EC ||
         (cfg->surface != NEMO_SURF_HLLC && cfg->surface != NEMO_SURF_RUSANOV)))
        return fail(NEMO_ERR_ARG, "the learned indicator is built for Euler, n_poly = 3, EC volume flux, HLLC or Rusanov");
    if (cfg->timedisc < 0 || cfg->timedisc > 1) return fail(NEMO_ERR_ARG, "invalid timedisc");
    if (cfg->precision < 0 || cfg->precision > 1) return fail(NEMO_ERR_ARG, "precision must be 0 (FP64) or 1 (FP32 operator)");
    if (!(cfg->gamma > 1.0)) return fail(NEMO_ERR_ARG, "gamma must exceed 1");
    Solver* s = new Solver();
    s->cfg = *cfg;
    s->M = cfg->n_poly + 1;
    s->M3 = s->M * s->M * s->M;
    s->V = cfg->equations == NEMO_EQ_GLM_MHD ? 9 : 5;
    build_host_ops(cfg->n_poly, s->ops);
    s->rk = make_rk();
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return fail(NEMO_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (hipMalloc(&s->dt_dev, 8 * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->mlp_p, nemo::MLP_NP * sizeof(float)) != hipSuccess ||
        hipMalloc(&s->mlp_inf, nemo::MLP_NP * sizeof(float)) != hipSuccess ||
        hipMalloc(&s->mlp_m, nemo::MLP_NP * sizeof(float)) != hipSuccess ||
        hipMalloc(&s->mlp_v, nemo::MLP_NP * sizeof(float)) != hipSuccess ||
        hipMalloc(&s->mlp_stats, 3 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&s->status, sizeof(int)) != hipSuccess ||
        hipMalloc(&s->cfl_red, 3 * nemo::CFL_GRID * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->cfl_ticket, sizeof(unsigned int)) != hipSuccess) {
        nemo_gpu_destroy(s);  // frees whichever of them were allocated
        return fail(NEMO_ERR_HIP, "hipMalloc failed");
    }
    // this device's copy of the allocation-free analysis scratch
    if (hipError_t e = scratch_device_ptr(&s->scratch); e != hipSuccess) {
        s->scratch = nullptr;
        nemo_gpu_destroy(s);
        return fail(NEMO_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    (void)hipMemset(s->status, 0, sizeof(int));
    (void)hipMemset(s->cfl_ticket, 0, sizeof(unsigned int));
    (void)hipMemset(s->dt_dev, 0, 8 * sizeof(double));
    (void)hipMemset(s->mlp_stats, 0, 3 * sizeof(unsigned long long));
    *handle = s;
    return NEMO_OK;
}

int nemo_gpu_destroy(void* handle) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return NEMO_OK;
    (void)hipStreamSynchronize(s->stream);
    free_mesh_buffers(*s);
    (void)hipFree(s->dt_dev);
    (void)hipFree(s->status);
    (void)hipFree(s->counter);
    (void)hipFree(s->units);
    (void)hipFree(s->an_part);
    (void)hipFree(s->an_out);
    (void)hipFree(s->mlp_p);
    (void)hipFree(s->mlp_inf);
    (void)hipFree(s->mlp_m);
    (void)hipFree(s->mlp_v);
    (void)hipFree(s->mlp_stats);
    (void)hipFree(s->hist);
    (void)hipFree(s->cfl_red);
    (void)hipFree(s->cfl_ticket);
    for (auto& p : s->ev) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    (void)hipStreamDestroy(s->stream);
    delete s;
    return NEMO_OK;
}

int nemo_gpu_upload_mesh(void* handle, int64_t nelem, const int32_t* nbr, const double* delta) {
    return nemo_gpu_upload_mesh_amr(handle, nelem, nbr, delta, 0, nullptr);
}

int nemo_gpu_upload_mesh_amr(void* handle, int64_t nelem, const int32_t* nbr, const double* delta, int64_t nmortar,
                             const int32_t* mortars) {
    return nemo_gpu_upload_mesh_bc(handle, nelem, nbr, delta, nmortar, mortars, 0, nullptr);
}

int nemo_gpu_upload_mesh_bc(void* handle, int64_t nelem, const int32_t* nbr, const double* delta, int64_t nmortar,
                            const int32_t* mortars, int64_t nbound, const int32_t* bfaces) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    if (int rc = check_mesh(*s, nelem, nbr, delta, nmortar, mortars, nbound, bfaces)) return rc;
    return set_mesh(*s, nelem, nbr, delta, nmortar, mortars, nbound, bfaces);
}

int nemo_gpu_remesh(void* handle, int64_t nelem, const int32_t* nbr, const double* delta, int64_t nmortar,
                    const int32_t* mortars, int64_t nunits, const int32_t* units, int64_t* fallbacks) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (int rc = check_mesh(*s, nelem, nbr, delta, nmortar, mortars)) return rc;
    if (nunits <= 0 || !units) return fail(NEMO_ERR_ARG, "invalid replacement units");
    // the units must pair every old element with every new one, in order
    long long io = 0, jn = 0;
    for (int64_t u = 0; u < nunits; ++u) {
        const int32_t* t = units + 3 * u;
        if (t[1] != io || t[2] != jn) return fail(NEMO_ERR_ARG, "replacement units out of order");
        switch (t[0]) {
        case nemo::UNIT_COPY: io += 1; jn += 1; break;
        case nemo::UNIT_REFINE: io += 1; jn += 8; break;
        case nemo::UNIT_COARSEN: io += 8; jn += 1; break;
        default: return fail(NEMO_ERR_ARG, "unknown replacement unit kind");
        }
        if (io > s->nelem || jn > nelem) return fail(NEMO_ERR_ARG, "replacement units exceed the meshes");
    }
    if (io != s->nelem || jn != nelem) return fail(NEMO_ERR_ARG, "replacement units do not cover both meshes");

    const size_t new_bytes = size_t(s->V) * size_t(nelem) * size_t(s->M3) * sizeof(double);
    // The new state goes to the idle ping-pong buffer when the new mesh fits
    // the buffers' capacity (no allocation); otherwise to a buffer of its own,
    // copied into the regrown state buffer once the mesh is set.
    const bool fits = nelem <= s->cap;
    DeviceBuffer u_new;  // freed on every early return
    if (!fits) HIP_TRY(hipMalloc(&u_new.p, new_bytes));
    if (nunits > s->units_cap) {
        (void)hipFree(s->units);
        s->units = nullptr;
        s->units_cap = 0;
        HIP_TRY(hipMalloc(&s->units, size_t(nunits) * 3 * sizeof(int32_t)));
        s->units_cap = nunits;
    }
    HIP_TRY(hipMemcpyAsync(s->units, units, size_t(nunits) * 3 * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->counter, 0, sizeof(int), s->stream));
    double* un = fits ? s->u[s->cur ^ 1] : static_cast<double*>(u_new.p);
    const int32_t* du = s->units;
    switch (s->M) {
    case 3: launch_remap<3>(*s, un, du, nunits, nelem); break;
    case 4: launch_remap<4>(*s, un, du, nunits, nelem); break;
    case 5: launch_remap<5>(*s, un, du, nunits, nelem); break;
    default: return fail(NEMO_ERR_ARG, "n_poly must be 2, 3 or 4");
    }
    HIP_TRY(hipGetLastError());
    int nfb = 0;
    HIP_TRY(hipMemcpyAsync(&nfb, s->counter, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (fallbacks) *fallbacks = nfb;
    if (fits) {
        if (int rc = set_mesh(*s, nelem, nbr, delta, nmortar, mortars)) return rc;
        s->cur ^= 1;
    } else {
        // set_mesh regrows every mesh-sized buffer (the old state is dead now)
        if (int rc = set_mesh(*s, nelem, nbr, delta, nmortar, mortars)) return rc;
        HIP_TRY(hipMemcpyAsync(s->u[s->cur], un, new_bytes, hipMemcpyDeviceToDevice, s->stream));
    }
    HIP_TRY(hipMemsetAsync(s->r0, 0, new_bytes, s->stream));
    HIP_TRY(hipMemsetAsync(s->r1, 0, new_bytes, s->stream));
    HIP_TRY(hipMemsetAsync(s->alpha_step, 0, size_t(nelem) * sizeof(double), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->have_state = true;
    return NEMO_OK;
}

int nemo_gpu_upload_state(void* handle, const double* u) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !u) return fail(NEMO_ERR_ARG, "null argument");
    if (!s->have_mesh) return fail(NEMO_ERR_NOT_READY, "mesh not uploaded");
    std::vector<double> dev;
    to_device_layout(*s, u, dev);
    const size_t bytes = s->state_doubles() * sizeof(double);
    HIP_TRY(hipMemcpyAsync(s->u[s->cur], dev.data(), bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->r0, 0, bytes, s->stream));
    HIP_TRY(hipMemsetAsync(s->r1, 0, bytes, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->have_state = true;
    return NEMO_OK;
}

int nemo_gpu_download_state(void* handle, double* u) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (!u) return fail(NEMO_ERR_ARG, "null argument");
    std::vector<double> dev(s->state_doubles());
    HIP_TRY(hipMemcpyAsync(dev.data(), s->u[s->cur], dev.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    to_host_layout(*s, dev, u);
    return NEMO_OK;
}

int nemo_gpu_rhs(void* handle, double* rhs) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (!rhs) return fail(NEMO_ERR_ARG, "null argument");
    if (int rc = launch_stage(*s, 1, 0.0, nullptr, s->rhs, nullptr)) return rc;
    std::vector<double> dev(s->state_doubles());
    HIP_TRY(hipMemcpyAsync(dev.data(), s->rhs, dev.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    to_host_layout(*s, dev, rhs);
    return NEMO_OK;
}

int nemo_gpu_indicator(void* handle, double* alpha) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (!alpha) return fail(NEMO_ERR_ARG, "null argument");
    if (s->cfg.indicator == NEMO_IND_NONE) {
        for (long long e = 0; e < s->nelem; ++e) alpha[e] = 0.0;
        return NEMO_OK;
    }
    // the Euler stage kernel stops after the indicator (no operator, no RHS)
    LaunchOpts o;
    o.alpha_only = s->cfg.equations == NEMO_EQ_EULER;
    if (int rc = launch_stage(*s, 1, 0.0, nullptr, s->rhs, s->alpha, nullptr, o)) return rc;
    HIP_TRY(hipMemcpyAsync(alpha, s->alpha, size_t(s->nelem) * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return NEMO_OK;
}

int nemo_gpu_analyze(void* handle, double* sums) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (!sums) return fail(NEMO_ERR_ARG, "null argument");
    const long long nodes = s->nelem * (long long)s->M3;
    const long long nblocks = (nodes + nemo::AN_NT - 1) / nemo::AN_NT;
    if (nblocks > s->an_cap) {
        (void)hipFree(s->an_part);
        s->an_part = nullptr;
        HIP_TRY(hipMalloc(&s->an_part, size_t(nblocks) * nemo::AN_NSUMS * sizeof(double)));
        s->an_cap = nblocks;
    }
    if (!s->an_out) HIP_TRY(hipMalloc(&s->an_out, nemo::AN_NSUMS * sizeof(double)));
    const bool mhd = s->cfg.equations == NEMO_EQ_GLM_MHD;
    switch (s->M) {
    case 3: launch_analyze<3>(*s, nblocks, mhd, s->an_part); break;
    case 4: launch_analyze<4>(*s, nblocks, mhd, s->an_part); break;
    case 5: launch_analyze<5>(*s, nblocks, mhd, s->an_part); break;
    default: return fail(NEMO_ERR_ARG, "n_poly must be 2, 3 or 4");
    }
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(nemo::analyze_finalize, dim3(1), dim3(nemo::AN_NT), 0, s->stream, s->an_part, nblocks, s->an_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(sums, s->an_out, nemo::AN_NSUMS * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return NEMO_OK;
}

int nemo_gpu_analyze_noalloc(void* handle, double* sums_dev) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (!sums_dev) return fail(NEMO_ERR_ARG, "null argument");
    const long long nodes = s->nelem * (long long)s->M3;
    if (nodes > NEMO_ANALYZE_NOALLOC_MAX_NODES)
        return fail(NEMO_ERR_ARG, "nemo_gpu_analyze_noalloc: " + std::to_string(nodes) +
                                      " nodes per variable exceed its scratch of " +
                                      std::to_string((long long)NEMO_ANALYZE_NOALLOC_MAX_NODES) +
                                      " (use nemo_gpu_analyze)");
    if (!s->scratch) return fail(NEMO_ERR_NOT_READY, "scratch address not resolved");
    const long long nblocks = (nodes + nemo::AN_NT - 1) / nemo::AN_NT;
    const bool mhd = s->cfg.equations == NEMO_EQ_GLM_MHD;
    // the same kernels as nemo_gpu_analyze; every partial row is written
    // before it is read, so nothing needs clearing
    switch (s->M) {
    case 3: launch_analyze<3>(*s, nblocks, mhd, s->scratch); break;
    case 4: launch_analyze<4>(*s, nblocks, mhd, s->scratch); break;
    case 5: launch_analyze<5>(*s, nblocks, mhd, s->scratch); break;
    default: return fail(NEMO_ERR_ARG, "n_poly must be 2, 3 or 4");
    }
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(nemo::analyze_finalize, dim3(1), dim3(nemo::AN_NT), 0, s->stream, s->scratch, nblocks, sums_dev);
    HIP_TRY(hipGetLastError());
    return NEMO_OK;
}

int nemo_gpu_calc_timestep(void* handle, double* dt, double* lmax, double* vmax) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (int rc = launch_cfl(*s, -1)) return rc;
    double h[3];
    HIP_TRY(hipMemcpyAsync(h, s->dt_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (dt) *dt = h[0];
    if (lmax) *lmax = h[1];
    if (vmax) *vmax = h[2];
    return NEMO_OK;
}

int nemo_gpu_set_glm(void* handle, double ch, double alpha) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    if (!(ch >= 0.0) || !(alpha >= 0.0)) return fail(NEMO_ERR_ARG, "GLM speed and damping must be >= 0");
    const double v[2] = {ch, alpha};
    HIP_TRY(hipMemcpyAsync(s->dt_dev + 3, v, sizeof(v), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return NEMO_OK;
}

int nemo_gpu_get_glm(void* handle, double* ch, double* alpha) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    double v[2];
    HIP_TRY(hipMemcpyAsync(v, s->dt_dev + 3, sizeof(v), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (ch) *ch = v[0];
    if (alpha) *alpha = v[1];
    return NEMO_OK;
}

int nemo_gpu_stage(void* handle, int32_t stage, double dt) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (stage < 1 || stage > 5) return fail(NEMO_ERR_ARG, "stage must be 1..5");
    LaunchOpts o;
    o.refresh = stage == 1;
    if (int rc = launch_stage(*s, stage, dt, nullptr, nullptr, nullptr, nullptr, o)) return rc;
    if (stage == 1) ++s->step_counter;
    s->cur ^= 1;
    return NEMO_OK;
}

int nemo_gpu_step(void* handle, double dt) {
    for (int st = 1; st <= 5; ++st)
        if (int rc = nemo_gpu_stage(handle, st, dt)) return rc;
    return NEMO_OK;
}

int nemo_gpu_run(void* handle, int32_t nsteps, double* dt_hist) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (nsteps < 0) return fail(NEMO_ERR_ARG, "nsteps must be >= 0");
    if (dt_hist && s->hist_cap < nsteps) {
        (void)hipFree(s->hist);
        s->hist = nullptr;
        HIP_TRY(hipMalloc(&s->hist, size_t(nsteps) * sizeof(double)));
        s->hist_cap = nsteps;
    }
    // The first step's dt comes from a standalone reduction; every later dt
    // from the CFL partials that the last stage of the previous step writes.
    if (nsteps > 0)
        if (int rc = launch_cfl(*s, dt_hist ? 0 : -1)) return rc;
    for (int it = 0; it < nsteps; ++it) {
        const bool more = it + 1 < nsteps;
        for (int st = 1; st <= 5; ++st) {
            double* part = (st == 5 && more) ? s->part : nullptr;
            LaunchOpts o;
            o.refresh = st == 1;
            if (int rc = launch_stage(*s, st, 0.0, s->dt_dev, nullptr, nullptr, part, o)) return rc;
            if (st == 1) ++s->step_counter;
            s->cur ^= 1;
        }
        if (more)
            if (int rc = launch_finalize(*s, dt_hist ? it + 1 : -1)) return rc;
    }
    if (dt_hist && nsteps > 0)
        HIP_TRY(hipMemcpyAsync(dt_hist, s->hist, size_t(nsteps) * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return status_check(s);
}

int nemo_gpu_synchronize(void* handle) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return status_check(s);
}

int nemo_gpu_get_status(void* handle, int32_t* code) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !code) return fail(NEMO_ERR_ARG, "null argument");
    int c = 0;
    HIP_TRY(hipMemcpyAsync(&c, s->status, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *code = c;
    return NEMO_OK;
}

// ---- learned indicator ------------------------------------------------------

int nemo_gpu_indicator_features(void* handle, float* features, double* alpha_pp) {
    Solver* s = static_cast<Solver*>(handle);
    if (int rc = check_ready(s)) return rc;
    if (!features) return fail(NEMO_ERR_ARG, "null argument");
    if (s->cfg.equations != NEMO_EQ_EULER || s->M != 4)
        return fail(NEMO_ERR_ARG, "indicator features are built for Euler, n_poly = 3");
    float* dfeat = nullptr;
    HIP_TRY(hipMalloc(&dfeat, size_t(s->nelem) * 4 * sizeof(float)));
    LaunchOpts o;
    o.feat_out = dfeat;
    int rc = launch_stage(*s, 1, 0.0, nullptr, s->rhs, s->alpha, nullptr, o);
    if (rc == NEMO_OK) {
        hipError_t e = hipMemcpyAsync(features, dfeat, size_t(s->nelem) * 4 * sizeof(float), hipMemcpyDeviceToHost,
                                      s->stream);
        if (e == hipSuccess && alpha_pp)
            e = hipMemcpyAsync(alpha_pp, s->alpha, size_t(s->nelem) * sizeof(double), hipMemcpyDeviceToHost,
                               s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(NEMO_ERR_HIP, std::string("indicator features: ") + hipGetErrorString(e));
    }
    (void)hipFree(dfeat);
    return rc;
}

int nemo_gpu_mlp_set_params(void* handle, const float* params) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !params) return fail(NEMO_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(s->mlp_p, params, nemo::MLP_NP * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->mlp_m, 0, nemo::MLP_NP * sizeof(float), s->stream));
    HIP_TRY(hipMemsetAsync(s->mlp_v, 0, nemo::MLP_NP * sizeof(float), s->stream));
    hipLaunchKernelGGL(nemo::mlp_inference_layout_kernel, dim3((nemo::MLP_NP + 255) / 256), dim3(256), 0, s->stream,
                       s->mlp_p, s->mlp_inf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->mlp_ready = true;
    return NEMO_OK;
}

int nemo_gpu_mlp_get_params(void* handle, float* params) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !params) return fail(NEMO_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(params, s->mlp_p, nemo::MLP_NP * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return NEMO_OK;
}

int nemo_gpu_mlp_set_policy(void* handle, double sample_fraction, double tolerance) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    if (!(sample_fraction >= 0.0 && sample_fraction <= 1.0) || !(tolerance >= 0.0))
        return fail(NEMO_ERR_ARG, "sample fraction must be in [0, 1] and the tolerance >= 0");
    s->sample_fraction = sample_fraction;
    s->judge_tol = tolerance;
    return NEMO_OK;
}

int nemo_gpu_mlp_stats(void* handle, int64_t* judged, int64_t* overridden, int64_t* fallback) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    unsigned long long h[3];
    HIP_TRY(hipMemcpyAsync(h, s->mlp_stats, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemsetAsync(s->mlp_stats, 0, sizeof(h), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (judged) *judged = int64_t(h[0]);
    if (overridden) *overridden = int64_t(h[1]);
    if (fallback) *fallback = int64_t(h[2]);
    return NEMO_OK;
}

int nemo_gpu_mlp_forward(void* handle, int64_t n, const float* features, float* y) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !features || !y || n <= 0) return fail(NEMO_ERR_ARG, "invalid arguments");
    if (!s->mlp_ready) return fail(NEMO_ERR_NOT_READY, "no network parameters");
    float *df = nullptr, *dy = nullptr;
    HIP_TRY(hipMalloc(&df, size_t(n) * 4 * sizeof(float)));
    if (hipMalloc(&dy, size_t(n) * sizeof(float)) != hipSuccess) {
        (void)hipFree(df);
        return fail(NEMO_ERR_HIP, "hipMalloc failed");
    }
    hipError_t e = hipMemcpyAsync(df, features, size_t(n) * 4 * sizeof(float), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        const long long groups = (n + 3) / 4;
        hipLaunchKernelGGL(nemo::mlp_forward_kernel, dim3(unsigned(groups < 65536 ? groups : 65536)), dim3(64), 0,
                           s->stream, s->mlp_inf, (long long)n, df, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, size_t(n) * sizeof(float), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(df);
    (void)hipFree(dy);
    if (e != hipSuccess) return fail(NEMO_ERR_HIP, std::string("mlp forward: ") + hipGetErrorString(e));
    return NEMO_OK;
}

namespace {

// One minibatch: partial gradients per block, then the ordered reduction
// (and an Adam step when lr > 0).
int mlp_batch(Solver* s, const float* df, const float* dt, const long long* didx, long long off, int B, float* part,
              float* part_loss, float* dgrad, float* dloss, int nparts, float lr, long long t) {
    hipLaunchKernelGGL(nemo::mlp_grad_kernel, dim3(nparts), dim3(64), 0, s->stream, s->mlp_p, df, dt, didx, off, B,
                       part, part_loss);
    HIP_TRY(hipGetLastError());
    const float c1 = float(1.0 - std::pow(0.9, double(t))), c2 = float(1.0 - std::pow(0.999, double(t)));
    hipLaunchKernelGGL(nemo::mlp_reduce_kernel, dim3((nemo::MLP_NP + 255) / 256), dim3(256), 0, s->stream, part,
                       nparts, part_loss, dgrad, dloss, s->mlp_p, s->mlp_m, s->mlp_v, lr > 0.0f ? 1 : 0, lr, c1, c2);
    HIP_TRY(hipGetLastError());
    return NEMO_OK;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
};

}  // namespace

int nemo_gpu_mlp_grad(void* handle, int64_t n, const float* features, const float* target, float* loss,
                      float* grad) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !features || !target || n <= 0 || n > (1 << 30)) return fail(NEMO_ERR_ARG, "invalid arguments");
    const int nparts = int(n < 128 ? n : 128);
    DevBuf f, t, part, pl, g, l;
    HIP_TRY(hipMalloc(&f.p, size_t(n) * 4 * sizeof(float)));
    HIP_TRY(hipMalloc(&t.p, size_t(n) * sizeof(float)));
    HIP_TRY(hipMalloc(&part.p, size_t(nparts) * nemo::MLP_NP * sizeof(float)));
    HIP_TRY(hipMalloc(&pl.p, size_t(nparts) * sizeof(float)));
    HIP_TRY(hipMalloc(&g.p, nemo::MLP_NP * sizeof(float)));
    HIP_TRY(hipMalloc(&l.p, sizeof(float)));
    HIP_TRY(hipMemcpyAsync(f.p, features, size_t(n) * 4 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(t.p, target, size_t(n) * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(l.p, 0, sizeof(float), s->stream));
    if (int rc = mlp_batch(s, static_cast<float*>(f.p), static_cast<float*>(t.p), nullptr, 0, int(n),
                           static_cast<float*>(part.p), static_cast<float*>(pl.p), static_cast<float*>(g.p),
                           static_cast<float*>(l.p), nparts, 0.0f, 1))
        return rc;
    if (grad) HIP_TRY(hipMemcpyAsync(grad, g.p, nemo::MLP_NP * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (loss) HIP_TRY(hipMemcpyAsync(loss, l.p, sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return NEMO_OK;
}

int nemo_gpu_mlp_train(void* handle, int64_t n, const float* features, const float* target, int32_t epochs,
                       int32_t batch, float lr, uint64_t seed, float* loss_per_epoch) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !features || !target || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0f))
        return fail(NEMO_ERR_ARG, "invalid arguments");
    if (!s->mlp_ready) return fail(NEMO_ERR_NOT_READY, "set initial parameters first (nemo_gpu_mlp_set_params)");
    const int nparts = batch < 128 ? batch : 128;
    DevBuf f, t, idx, part, pl, l;
    HIP_TRY(hipMalloc(&f.p, size_t(n) * 4 * sizeof(float)));
    HIP_TRY(hipMalloc(&t.p, size_t(n) * sizeof(float)));
    HIP_TRY(hipMalloc(&idx.p, size_t(n) * sizeof(long long)));
    HIP_TRY(hipMalloc(&part.p, size_t(nparts) * nemo::MLP_NP * sizeof(float)));
    HIP_TRY(hipMalloc(&pl.p, size_t(nparts) * sizeof(float)));
    HIP_TRY(hipMalloc(&l.p, size_t(epochs) * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(f.p, features, size_t(n) * 4 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(t.p, target, size_t(n) * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(l.p, 0, size_t(epochs) * sizeof(float), s->stream));
    std::vector<long long> perm(static_cast<size_t>(n));
    for (long long i = 0; i < n; ++i) perm[size_t(i)] = i;
    uint64_t state = seed ^ 0x9E3779B97F4A7C15ull;
    auto next = [&state]() {  // SplitMix64
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    long long t_step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (long long i = n - 1; i > 0; --i) std::swap(perm[size_t(i)], perm[size_t(next() % uint64_t(i + 1))]);
        HIP_TRY(hipMemcpyAsync(idx.p, perm.data(), size_t(n) * sizeof(long long), hipMemcpyHostToDevice, s->stream));
        for (long long off = 0; off + batch <= n; off += batch) {
            ++t_step;
            if (int rc = mlp_batch(s, static_cast<float*>(f.p), static_cast<float*>(t.p),
                                   static_cast<long long*>(idx.p), off, batch, static_cast<float*>(part.p),
                                   static_cast<float*>(pl.p), nullptr, static_cast<float*>(l.p) + ep, nparts, lr,
                                   t_step))
                return rc;
        }
        // the host permutation buffer is rewritten next epoch: wait for this one
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    if (loss_per_epoch) {
        std::vector<float> h(static_cast<size_t>(epochs));
        HIP_TRY(hipMemcpyAsync(h.data(), l.p, size_t(epochs) * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        const long long nb = n / batch;
        for (int ep = 0; ep < epochs; ++ep) loss_per_epoch[ep] = nb > 0 ? h[size_t(ep)] / float(nb) : 0.0f;
    }
    hipLaunchKernelGGL(nemo::mlp_inference_layout_kernel, dim3((nemo::MLP_NP + 255) / 256), dim3(256), 0, s->stream,
                       s->mlp_p, s->mlp_inf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    return NEMO_OK;
}

// Plain copy of n doubles (the known-size kernel that calibrates the HBM
// traffic counters: it reads and writes exactly 8n bytes).
__global__ void nemo_copy_kernel(const double* __restrict__ src, double* __restrict__ dst, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int nemo_gpu_copy_calibration(void* handle, int64_t nbytes) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || nbytes <= 0 || nbytes % 8) return fail(NEMO_ERR_ARG, "nbytes must be a positive multiple of 8");
    double *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, size_t(nbytes)));
    if (hipMalloc(&b, size_t(nbytes)) != hipSuccess) {
        (void)hipFree(a);
        return fail(NEMO_ERR_HIP, "hipMalloc failed");
    }
    const long long n = nbytes / 8;
    hipError_t e = hipMemsetAsync(a, 0, size_t(nbytes), s->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(nemo_copy_kernel, dim3(4096), dim3(256), 0, s->stream, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(NEMO_ERR_HIP, std::string("copy calibration: ") + hipGetErrorString(e));
    return NEMO_OK;
}

// Arithmetic self-check: n operand pairs from a counter-based generator
// (a with a random exponent in [-60, 60] and mantissa; b independent for odd
// i, a (1 + d) with |d| < 10^-k, k = 1..16, for even i), comparing the bits
// of div_rn(a, b) with the IEEE quotient and of rcp_nr(b) with the IEEE
// reciprocal.
namespace {
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double random_double(unsigned long long h) {
    const long long ex = 1023 + (long long)((h >> 52) % 121) - 60;
    return __longlong_as_double((ex << 52) | (long long)(h & 0xFFFFFFFFFFFFFull));
}
__global__ void nemo_arith_check_kernel(long long n, unsigned long long seed, unsigned long long* counts) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const unsigned long long h1 = mix64(seed ^ (2ull * i)), h2 = mix64(seed ^ (2ull * i + 1));
        const double a = random_double(h1);
        double b;
        if (i & 1) {
            b = random_double(h2);
        } else {
            double scale = 0.1;
            for (int k = 0; k < int(h2 & 15); ++k) scale *= 0.1;
            b = a * (1.0 + (double(h2 >> 11) * 0x1.0p-53 - 0.5) * scale);
        }
        if (__double_as_longlong(nemo::div_rn(a, b)) != __double_as_longlong(a / b)) atomicAdd(&counts[0], 1ull);
        if (__double_as_longlong(nemo::rcp_nr(b)) != __double_as_longlong(1.0 / b)) atomicAdd(&counts[1], 1ull);
    }
}
}  // namespace

int nemo_gpu_check_arithmetic(void* handle, int64_t n, uint64_t seed, int64_t* div_mismatches,
                              int64_t* rcp_mismatches) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || n <= 0 || !div_mismatches || !rcp_mismatches) return fail(NEMO_ERR_ARG, "bad argument");
    unsigned long long* d = nullptr;
    HIP_TRY(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(nemo_arith_check_kernel, dim3(2048), dim3(256), 0, s->stream, (long long)n,
                           (unsigned long long)seed, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(NEMO_ERR_HIP, std::string("arithmetic check: ") + hipGetErrorString(e));
    *div_mismatches = int64_t(h[0]);
    *rcp_mismatches = int64_t(h[1]);
    return NEMO_OK;
}

int nemo_gpu_set_timing(void* handle, int32_t enable) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail(NEMO_ERR_ARG, "null handle");
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->timing = enable != 0;
    s->ev_used = 0;
    return NEMO_OK;
}

int nemo_gpu_read_timing(void* handle, int64_t* launches, double* total_ms) {
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !launches || !total_ms) return fail(NEMO_ERR_ARG, "null argument");
    HIP_TRY(hipStreamSynchronize(s->stream));
    double tot = 0.0;
    for (size_t i = 0; i < s->ev_used; ++i) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev[i].first, s->ev[i].second));
        tot += double(ms);
    }
    *launches = int64_t(s->ev_used);
    *total_ms = tot;
    s->ev_used = 0;
    return NEMO_OK;
}

}  // extern "C"
