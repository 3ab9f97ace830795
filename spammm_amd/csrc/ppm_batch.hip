// ppm_batch.hip — batched forward: many molecules (uneven atom counts, cell sizes and grid
// shapes) through grid projection -> relaxed scan -> df in ONE persistent launch.
//
// Work items: for each molecule m its grid tiles (256 columns x vz planes of the stored z window)
// and its relaxation blocks (4 wavefronts x 64 pixels; each lane relaxes one pixel through all
// z-steps and then applies the df operator to its own Fz column), as two queue segments; the host
// orders the segments (by default each molecule's blocks right after its tiles; with a lag k the
// tiles of m + k come before the blocks of m, so a block rarely waits for its tiles).
// Persistent workgroups take items from an atomic counter.  A relaxation block of m waits (with
// a bounded spin) until every grid tile of m has been published: the tile's stores are released
// at agent scope before the per-molecule counter is incremented, and the waiting workgroup
// acquires at agent scope before it reads the grid, so the data crosses XCD L2 caches correctly.
// Items are taken in increasing order, so a waited-for tile has always been taken by a running
// workgroup: the wait ends.  Every workgroup leaves the loop when the queue is exhausted.
//
// The grid and relaxation arithmetic is the device code of the plain kernels
// (ppm_grid_core.hpp, ppm_relax_core.hpp), so each molecule's results equal the plain path's
// on the same inputs, bit for bit, whenever no probe leaves the stored z window (a probe that
// does sets the molecule's flag; the caller recomputes that molecule on the plain path).
#include <mutex>

#include "ppm_grid_core.hpp"
#include "ppm_relax_core.hpp"

namespace ppm {
namespace {

using gc::GridArgs;
using gc::Site;


constexpr int kSpinLimit = 1 << 20;     // polls (~1 us each with the sleep): a bound of about a second

struct Queue {           // device workspace, zeroed by k_batch_reset
    int head;            // next work item
    int error;           // bit 1: a wait hit the spin limit
    int pad[2];
    // followed by: int done[nmol]; int flag[nmol];
};

__device__ __forceinline__ int* q_done(Queue* q) { return reinterpret_cast<int*>(q + 1); }
__device__ __forceinline__ int* q_flag(Queue* q, int nmol) { return reinterpret_cast<int*>(q + 1) + nmol; }

__global__ void k_batch_reset(Queue* q, int nmol, int head_only) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { q->head = 0; if (!head_only) q->error = 0; }
    if (!head_only && i < 2 * nmol) q_done(q)[i] = 0;
}

// status[0] = error word, status[1 + m] = window flag of molecule m (the allocation-free launch's
// status, copied out of its scratch queue).
__global__ void k_batch_status(Queue* q, int nmol, int* __restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) status[0] = q->error;
    if (i < nmol) status[1 + i] = q_flag(q, nmol)[i];
}

__device__ GridArgs mol_grid_args(const PpmMolDesc& d, const PpmBatchParams& P) {
    GridArgs a{};
    a.nx = d.nx; a.ny = d.ny; a.nz = d.nz; a.ncol = d.nx * d.ny;
    for (int k = 0; k < 3; ++k) {
        a.p0[k] = d.p0[k]; a.dA[k] = d.dA[k]; a.dB[k] = d.dB[k]; a.dC[k] = d.dC[k];
        a.lvA[k] = d.lvA[k]; a.lvB[k] = d.lvB[k];
    }
    a.na = d.na;
    a.npbc_a = d.npbc_a; a.npbc_b = d.npbc_b;
    a.nimg_a = 2 * d.npbc_a + 1; a.nimg_b = 2 * d.npbc_b + 1;
    a.nsites = d.na * a.nimg_a * a.nimg_b;
    a.cap = P.cap;
    a.q0 = 0.f;
    a.noff = 0;
    // compaction with static indices only (a dynamically indexed private array lives in scratch)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (P.Qs[k] == 0.f) continue;
        if (P.QZs[k] == 0.f) {
            a.q0 += P.Qs[k] * kCOULOMB;
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t == a.noff) { a.qoff[t] = P.Qs[k] * kCOULOMB; a.zoff[t] = P.QZs[k]; }
            ++a.noff;
        }
    }
    return a;
}

// First queue item of a molecule's relaxation blocks (right after its tiles when relax_item_off = 0).
__device__ __forceinline__ int relax_start(const PpmMolDesc& d) {
    return d.relax_item_off > 0 ? d.relax_item_off : d.item_off + d.n_grid_items;
}

__device__ rc::TiltArgs tilt_args(const PpmBatchParams& P) {
    rc::TiltArgs t{};
    t.nz = P.nz; t.max_iter = P.ctl.max_iter; t.check_before = P.ctl.check_before; t.f2conv = P.ctl.f2conv;
    t.damped = P.ctl.damped != 0;
    for (int q = 0; q < 3; ++q) { t.A[q] = P.tip.tipA[q]; t.B[q] = P.tip.tipB[q]; t.C[q] = P.tip.tipC[q]; }
    t.dtip = P.tip.tipC[3];
    for (int q = 0; q < 4; ++q) { t.k[q] = P.tip.stiffness[q]; t.d0[q] = P.tip.dpos0[q]; }
    t.dt0 = P.tip.relax[0]; t.damp0 = P.tip.relax[1]; t.surf = P.tip.surfFF[0];
    return t;
}

template <int MODEL, bool HASQ0, int V, bool OFFQ>
__device__ void do_grid_item(const PpmMolDesc& d, const PpmBatchParams& P, int local, const float4* atoms,
                             const float* cpar, float4* grids, Site* sites) {
    const GridArgs a = mol_grid_args(d, P);
    const int ncolblk = (a.ncol + 255) / 256;
    const int cb = local % ncolblk, zb = local / ncolblk;
    const int cs = MODEL == PPM_MODEL_LJ ? 2 : 4;
    gc::grid_tile<MODEL, HASQ0, V, true, OFFQ ? gc::kOffStatic : gc::kOffNone>(atoms + d.atom_off, cpar + (size_t)d.atom_off * cs, a, sites,
                                         cb * 256 + threadIdx.x, d.klo + zb * V, d.klo + d.nzw, grids + d.grid_off,
                                         d.klo);
}

__device__ void do_relax_item(const PpmMolDesc& d, const PpmBatchParams& P, int local, const float4* grids,
                              float4* FEs, float4* PPs, float* df, const float* dfW, const int* dfStart,
                              int* flag) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tw = P.tile_w, th = 64 / P.tile_w;
    const int tiles_y = (d.scan_ny + tw - 1) / tw;
    const int wt = local * 4 + wave;
    const int tx = wt / tiles_y, ty = wt - tx * tiles_y;
    const int ix = tx * th + lane / tw, iy = ty * tw + lane % tw;
    if (ix >= d.scan_nx || iy >= d.scan_ny) return;
    rc::GridDev g;
    g.data = grids + d.grid_off;
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz;
    for (int q = 0; q < 4; ++q) { g.da[q] = d.dinvA[q]; g.db[q] = d.dinvB[q]; g.dc[q] = d.dinvC[q]; }
    g.klo = d.klo; g.nzw = d.nzw; g.flag = flag;
    const float fx = (float)ix, fy = (float)iy;
    const f3 start = mk3(d.scan_p0[0] + d.scan_da[0] * fx + d.scan_db[0] * fy,
                          d.scan_p0[1] + d.scan_da[1] * fx + d.scan_db[1] * fy,
                          d.scan_p0[2] + d.scan_da[2] * fx + d.scan_db[2] * fy);
    const size_t pix = (size_t)d.pix_off + (size_t)ix * d.scan_ny + iy;
    const size_t o = pix * P.nz;
    const rc::TiltArgs t = tilt_args(P);
    const rc::DfOut dfo{dfW, dfStart, P.df_taps, P.df_nz_out, false};
    float4* pp = PPs ? PPs + o : nullptr;
    float* dr = df ? df + pix * P.df_nz_out : nullptr;
    if (rc::fast_frame(t, g))           // the same value in every lane of the item
        rc::relax_tilted_lane<false, true>(g, start, t, FEs + o, pp, df ? &dfo : nullptr, dr);
    else
        rc::relax_tilted_lane<false, false>(g, start, t, FEs + o, pp, df ? &dfo : nullptr, dr);
}

// OFFQ: the tip has charge sites at a nonzero z offset (compiled out when it has none).
template <int MODEL, bool HASQ0, int V, bool OFFQ>
__global__ __launch_bounds__(256) void k_batch(const PpmMolDesc* __restrict__ mols, PpmBatchParams P,
                                               const float4* __restrict__ atoms, const float* __restrict__ cpar,
                                               float4* __restrict__ grids, float4* __restrict__ FEs,
                                               float4* __restrict__ PPs, float* __restrict__ df,
                                               const float* __restrict__ dfW, const int* __restrict__ dfStart,
                                               Queue* q) {
    __shared__ Site sites[gc::kTile + 1];
    __shared__ int s_item, s_mol, s_local;
    int* done = q_done(q);
    int* flags = q_flag(q, P.nmol);
    // Molecule whose grid tile this workgroup finished in the previous iteration (-1: none).
    // Thread 0 publishes it at the top of the next iteration, in the same thread-0 section
    // as the dequeue: a thread-0-only section at the END of the loop body would be merged by
    // the compiler with the one at the top into a divergent inner loop whose other lanes
    // re-read the stale item after the barrier and never leave it.
    int signal_m = -1;
    while (true) {
        if (threadIdx.x == 0) {
            if (signal_m >= 0) __hip_atomic_fetch_add(&done[signal_m], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int item = atomicAdd(&q->head, 1);
            int m = -1, local = 0;
            if (item < P.nitems) {
                int lo = 0, hi = P.nmol - 1;            // last molecule with item_off <= item
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (mols[mid].item_off <= item) lo = mid; else hi = mid - 1;
                }
                if (item < mols[lo].item_off + mols[lo].n_grid_items) {
                    m = lo;
                    local = item - mols[lo].item_off;
                } else {                                // a relaxation block: last molecule whose
                    lo = 0;                             // relaxation segment starts at or before item
                    hi = P.nmol - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (relax_start(mols[mid]) <= item) lo = mid; else hi = mid - 1;
                    }
                    m = lo;
                    local = mols[lo].n_grid_items + (item - relax_start(mols[lo]));
                }
            }
            s_item = item;
            s_mol = m;
            s_local = local;
        }
        __syncthreads();
        const int item = __builtin_amdgcn_readfirstlane(s_item), m = __builtin_amdgcn_readfirstlane(s_mol);
        const int local = __builtin_amdgcn_readfirstlane(s_local);
        __syncthreads();
        if (item >= P.nitems) break;
        const PpmMolDesc& d = mols[m];
        signal_m = -1;
        if (local < d.n_grid_items) {
            if (P.mode != PPM_BATCH_RELAX_ONLY) {
                do_grid_item<MODEL, HASQ0, V, OFFQ>(d, P, local, atoms, cpar, grids, sites);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __syncthreads();            // every thread's tile stores precede the signal
                signal_m = m;
            }
        } else if (P.mode != PPM_BATCH_GRID_ONLY) {
            if (threadIdx.x == 0) {
                // relaxed polling (no cache invalidation per poll), one acquire after the wait
                int spins = 0;
                while (__hip_atomic_load(&done[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < d.n_grid_items) {
                    __builtin_amdgcn_s_sleep(8);
                    if (++spins > kSpinLimit) {
                        atomicOr(&q->error, 1);
                        break;
                    }
                }
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            do_relax_item(d, P, local - d.n_grid_items, grids, FEs, PPs, df, dfW, dfStart, &flags[m]);
        }
    }
}

template <int MODEL, bool HASQ0, bool OFFQ>
void launch_v(int V, dim3 grd, hipStream_t st, const PpmMolDesc* mols, const PpmBatchParams& P, const float4* atoms,
              const float* cpar, float4* grids, float4* FEs, float4* PPs, float* df, const float* dfW,
              const int* dfStart, Queue* q) {
    if (V == 8)
        hipLaunchKernelGGL((k_batch<MODEL, HASQ0, 8, OFFQ>), grd, dim3(256), 0, st, mols, P, atoms, cpar, grids, FEs,
                           PPs, df, dfW, dfStart, q);
    else
        hipLaunchKernelGGL((k_batch<MODEL, HASQ0, 4, OFFQ>), grd, dim3(256), 0, st, mols, P, atoms, cpar, grids, FEs,
                           PPs, df, dfW, dfStart, q);
}

template <int MODEL>
void launch_q(bool hasq0, bool offq, int V, dim3 grd, hipStream_t st, const PpmMolDesc* mols,
              const PpmBatchParams& P, const float4* atoms, const float* cpar, float4* grids, float4* FEs,
              float4* PPs, float* df, const float* dfW, const int* dfStart, Queue* q) {
    if (hasq0) {
        if (offq) launch_v<MODEL, true, true>(V, grd, st, mols, P, atoms, cpar, grids, FEs, PPs, df, dfW, dfStart, q);
        else launch_v<MODEL, true, false>(V, grd, st, mols, P, atoms, cpar, grids, FEs, PPs, df, dfW, dfStart, q);
    } else {
        if (offq) launch_v<MODEL, false, true>(V, grd, st, mols, P, atoms, cpar, grids, FEs, PPs, df, dfW, dfStart, q);
        else launch_v<MODEL, false, false>(V, grd, st, mols, P, atoms, cpar, grids, FEs, PPs, df, dfW, dfStart, q);
    }
}

}  // namespace

size_t batch_workspace(int nmol) { return sizeof(Queue) + 2 * (size_t)nmol * sizeof(int); }

// Scratch of the allocation-free batch entry point: the work queue, per-molecule tile counters and
// window flags of up to kBatchNoallocMaxMol molecules, one copy per GPU for the life of the process.
// In host code `scratch` names a host-side stand-in; kernels get its device address from
// hipGetSymbolAddress, looked up once per device.
static __device__ __align__(256) unsigned char scratch[sizeof(Queue) + 2 * (size_t)kBatchNoallocMaxMol * sizeof(int)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first call per device).
int noalloc_scratch(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("batch_forward_noalloc: no current HIP device (or device id %d >= %d)", dev, kMaxDevices);
        return PPM_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("batch_forward_noalloc: hipGetSymbolAddress of the scratch array failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return PPM_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PPM_OK;
}
}  // namespace

int launch_batch(const PpmMolDesc* mols, const PpmBatchParams& P, const float* atoms, const float* cpar,
                 float* grids, float* FEs, float* PPs, float* df, const float* dfW, const int* dfStart,
                 void* workspace, int workers, int reset, hipStream_t stream) {
    Queue* q = reinterpret_cast<Queue*>(workspace);
    if (reset) {
        const int n = 2 * P.nmol + 1;
        hipLaunchKernelGGL(k_batch_reset, dim3((n + 255) / 256), dim3(256), 0, stream, q, P.nmol, reset == 2);
        if (int r = check_launch("batch_reset")) return r;
    }
    bool hasq0 = false, offq = false;
    for (int k = 0; k < 4; ++k) {
        hasq0 |= (P.Qs[k] != 0.f && P.QZs[k] == 0.f);
        offq |= (P.Qs[k] != 0.f && P.QZs[k] != 0.f);
    }
    const dim3 grd(workers);
    const float4* at = reinterpret_cast<const float4*>(atoms);
    float4* gr = reinterpret_cast<float4*>(grids);
    float4* fe = reinterpret_cast<float4*>(FEs);
    float4* pp = reinterpret_cast<float4*>(PPs);
    if (P.model == PPM_MODEL_LJ)
        launch_q<PPM_MODEL_LJ>(hasq0, offq, P.vz, grd, stream, mols, P, at, cpar, gr, fe, pp, df, dfW, dfStart, q);
    else
        launch_q<PPM_MODEL_MORSE>(hasq0, offq, P.vz, grd, stream, mols, P, at, cpar, gr, fe, pp, df, dfW, dfStart, q);
    return check_launch("batch");
}

int launch_batch_noalloc(const PpmMolDesc* mols, const PpmBatchParams& P, const float* atoms, const float* cpar,
                         float* grids, float* FEs, float* PPs, float* df, const float* dfW, const int* dfStart,
                         int* status, int workers, int reset, hipStream_t stream) {
    void* ws = nullptr;
    if (int r = noalloc_scratch(&ws)) return r;
    if (int r = launch_batch(mols, P, atoms, cpar, grids, FEs, PPs, df, dfW, dfStart, ws, workers, reset, stream))
        return r;
    hipLaunchKernelGGL(k_batch_status, dim3((P.nmol + 255) / 256), dim3(256), 0, stream, reinterpret_cast<Queue*>(ws),
                       P.nmol, status);
    return check_launch("batch_status");
}

}  // namespace ppm
