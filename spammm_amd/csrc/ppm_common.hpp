// ppm_common.hpp — shared definitions of the gfx950 probe-particle AFM kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "ppm_api.h"

namespace ppm {

// Physical constants of kernels/common.cl:28-29.
constexpr float kR2SAFE = 1e-4f;
constexpr float kCOULOMB = 14.3996448915f;
constexpr double kR2SAFE_D = 1e-4;            // the same constants for the FP64 projection
constexpr double kCOULOMB_D = 14.3996448915;

// Algorithmic FLOP per pair evaluation (grid projection roofline unit), counted on
// the reference pair laws with add/sub/mul = 1, FMA = 2, rcp/sqrt/rsqrt/exp = 1:
//   getLJ (kernels/AFM.cl:738-745) 27 + one getCoulombAFM site (kernels/AFM.cl:730-736) 25 = 52
//   getMorse (kernels/AFM.cl:747-757) 29 + one getCoulombAFM site 25 = 54
constexpr int kFlopPairLJ = 52;
constexpr int kFlopPairMorse = 54;

struct f3 {
    float x, y, z;
};
__host__ __device__ inline f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

// Launch helpers shared by the translation units.
void set_error(const char* fmt, ...);
int check_launch(const char* what);

// Compute units and LDS bytes per compute unit of the current device, looked up once per device
// (MI355X's 256 and 160 KiB if a query fails).
int device_cus();
int device_lds_per_cu();

// Per-translation-unit launchers (host side).
int launch_grid_project(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec& g,
                        const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                        const float lvB[3], int cap, const float* base, float* grid, bool acc64,
                        hipStream_t stream, int klo = 0, int nzw = 0);
// klo/nzw/flag: the grid buffer holds planes [klo, klo + nzw) only (nzw = 0: the whole grid); a probe
// that samples outside them sets *flag.
int launch_relax_tilted(const float* grid, const PpmGridSpec& g, const float* pts, int n_scan, int nz,
                        const PpmTip& tip, const PpmRelaxCtl& ctl, float* FEs, float* PPs, hipStream_t stream,
                        int klo = 0, int nzw = 0, int* flag = nullptr, const float* dfW = nullptr,
                        const int* dfStart = nullptr, int taps = 0, int nout = 0, float* df = nullptr);
int launch_relax_strokes(const float* grid, const PpmGridSpec& g, const float* pts, int n_scan, int nz,
                         const float dTip[4], const float stiffness[4], const float dpos0[4], const float relax[4],
                         const PpmRelaxCtl& ctl, float* FEs, float* disps, float* PPs, hipStream_t stream);
int launch_relax_adjoint(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                         const PpmTip& tip, const float* PPs, const float* dL_dFEs, float* grid_cot, float* lambdas,
                         hipStream_t stream);
int launch_relax_strokes_2d(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                            float k_lat, float dh, float dt, float damp, const PpmRelaxCtl& ctl, float* FEs,
                            float* disps, float* PPs, hipStream_t stream);
int launch_relax_state_diag(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                            const PpmTip& tip, const float* PPs, float* diag, hipStream_t stream);
int launch_relax_polish(const float* grid, const PpmGridSpec& s, const float* pts, int n_scan, int nz,
                        const PpmTip& tip, int steps, float* PPs, float* FEs, hipStream_t stream);
int launch_grid_project_f64(int model, const double* atoms, const double* cpar, int na, const PpmGridSpec& g,
                            const double Qs[4], const double QZs[4], int npbc_a, int npbc_b, const double lvA[3],
                            const double lvB[3], int cap, double* grid, hipStream_t stream);
int launch_relax_tilted_f64(const double* grid, const PpmGridSpec& g, const double* pts, int n_scan, int nz,
                            const PpmTip& tip, const PpmRelaxCtl& ctl, double* FEs, double* PPs,
                            hipStream_t stream);
size_t fdbm_spectrum_bytes(int nx, int ny, int nz);
int release_fft_plans();
int launch_fft_combine(int mode, int nx, int ny, int nz, float step, const float* a, const float* b, float scale,
                       float* out, void* workspace, hipStream_t stream);
int launch_pauli_scale(float* field, long long n, float A, float beta, float clip, hipStream_t stream);
int launch_pad_roll(const float* src, const int s[3], float* dst, const int n[3], const int o[3], const int r[3],
                    hipStream_t stream);
int launch_flip3(const float* src, float* dst, const int n[3], hipStream_t stream);
int launch_dispersion(const float* atoms, int na, const float origin[3], float step, const int n[3], float ra2,
                      int accumulate, float* out, hipStream_t stream);
int launch_compose_gradient(const float* e1, const float* e2, const float* e3, const int n[3], float step,
                            float* grid, hipStream_t stream);
int ewald_num_g(int n_harm);
size_t ewald_system_bytes(int n_layers, int n_harm);
size_t ewald_kept_count_offset(int n_layers, int n_harm);
size_t ewald_grid_workspace(int n_harm, int nz);
int launch_ewald_prepare(const float* ions, const int* layer_off, int n_layers, const double a[2], const double b[2],
                         int n_harm, double drop_tol, void* system, hipStream_t stream);
int launch_ewald_points(int mode, const float* pts, int np, const void* system, int n_layers, int n_harm,
                        const double a[2], const double b[2], float* phi, hipStream_t stream);
int launch_ewald_brute(const float* pts, int np, const float* ions, int ni, const double a[2], const double b[2],
                       int n_rep, bool square, float* phi, hipStream_t stream);
int launch_ewald_cluster(const float* pts, int np, const float* ions, int ni, float* phi, hipStream_t stream);
int launch_ewald_tip_grid(const void* system, int n_layers, int n_harm, const double a[2], const double b[2],
                          const PpmGridSpec& gs, const float Qs[4], const float QZs[4], int accumulate,
                          void* workspace, float* grid, hipStream_t stream);
size_t batch_workspace(int nmol);
// Largest batch of the allocation-free batch entry point (its file-scope scratch holds the queue,
// counters and flags of this many molecules: 512 KiB + 16 B per GPU).
constexpr int kBatchNoallocMaxMol = PPM_BATCH_NOALLOC_MAX_MOL;
int launch_batch_noalloc(const PpmMolDesc* mols, const PpmBatchParams& P, const float* atoms, const float* cpar,
                         float* grids, float* FEs, float* PPs, float* df, const float* dfW, const int* dfStart,
                         int* status, int workers, int reset, hipStream_t stream);
int launch_batch(const PpmMolDesc* mols, const PpmBatchParams& P, const float* atoms, const float* cpar,
                 float* grids, float* FEs, float* PPs, float* df, const float* dfW, const int* dfStart,
                 void* workspace, int workers, int reset, hipStream_t stream);
size_t grid_adjoint_workspace(int na, const PpmGridSpec& g);
int launch_grid_adjoint(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec& g,
                        const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                        const float lvB[3], int cap, const float* grid, const float* grid_cot, const float* base,
                        void* workspace, int params, double* grad, hipStream_t stream);
int launch_fe_in_strokes(const float* grid, const PpmGridSpec& g, const float* pts, int n_scan, int nz,
                         const PpmTip& tip, const float dTip[4], const float dpos0[4], float* FEs,
                         hipStream_t stream);
int launch_direct_relax(const float* atoms, const float* cms, int na, const float* pts, int n_scan, int nz,
                        const PpmTip& tip, const float Qs[4], const float QZs[4], const PpmRelaxCtl& ctl, float* FEs,
                        float* PPs, hipStream_t stream);
size_t direct_vjp_workspace(int na, int n_scan, int nz);
int launch_direct_vjp(const float* atoms, const float* cms, int na, const float* pts, int n_scan, int nz,
                      const PpmTip& tip, const float Qs[4], const float QZs[4], const float* PPs, const float* dL_dFEs,
                      float* lambdas, float* diag, void* workspace, double* grad, hipStream_t stream);
int launch_axis_operator(const float* in, long long so, long long sa, long long si, const float chw[4], int nch,
                         int n_outer, int n_axis_in, int n_inner, const float* W, const int* start, int taps,
                         int n_axis_out, float alpha, int accumulate, float* out, hipStream_t stream);

}  // namespace ppm
