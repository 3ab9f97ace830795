// ppm_api.hip — extern "C" entry points of libppm_hip.so (declared in include/ppm_api.h).
//
// Each entry point validates its arguments on the host, then enqueues the kernels on
// the caller's stream.  Nothing is allocated and nothing is synchronized.
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "ppm_common.hpp"

namespace ppm {

static thread_local char g_err[512] = "";

namespace {
constexpr int kMaxDevices = 64;
std::atomic<int> g_cus[kMaxDevices], g_lds[kMaxDevices];

int device_attr(std::atomic<int>* cache, hipDeviceAttribute_t attr, int fallback) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return fallback;
    }
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, attr, dev) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            n = fallback;
        }
        cache[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}
}  // namespace

int device_cus() { return device_attr(g_cus, hipDeviceAttributeMultiprocessorCount, 256); }
int device_lds_per_cu() {
    return device_attr(g_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 160 * 1024);
}

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
        return PPM_ERR_HIP;
    }
    return PPM_OK;
}

static bool finite4(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

static int check_spec(const char* fn, const PpmGridSpec* s) {
    if (!s) { set_error("%s: spec is NULL", fn); return PPM_ERR_INVALID_ARG; }
    if (s->nx < 1 || s->ny < 1 || s->nz < 1) {
        set_error("%s: invalid grid dims (%d,%d,%d)", fn, s->nx, s->ny, s->nz);
        return PPM_ERR_INVALID_ARG;
    }
    if ((long long)s->nx * s->ny > (1LL << 30) || (long long)s->nx * s->ny * s->nz > (1LL << 34)) {
        set_error("%s: grid (%d,%d,%d) too large", fn, s->nx, s->ny, s->nz);
        return PPM_ERR_INVALID_ARG;
    }
    if (!finite4(s->dinvA, 4) || !finite4(s->dinvB, 4) || !finite4(s->dinvC, 4) || !finite4(s->p0, 3) ||
        !finite4(s->dA, 3) || !finite4(s->dB, 3) || !finite4(s->dC, 3)) {
        set_error("%s: non-finite grid geometry", fn);
        return PPM_ERR_INVALID_ARG;
    }
    return PPM_OK;
}

static int check_scan(const char* fn, const float* pts, int n_scan, int nz, const float* FEs) {
    if (n_scan < 1 || nz < 1) { set_error("%s: invalid n_scan=%d nz=%d", fn, n_scan, nz); return PPM_ERR_INVALID_ARG; }
    if ((long long)n_scan * nz > (1LL << 31)) { set_error("%s: n_scan*nz too large", fn); return PPM_ERR_INVALID_ARG; }
    if (!pts || !FEs) { set_error("%s: NULL points/FEs", fn); return PPM_ERR_INVALID_ARG; }
    if ((reinterpret_cast<uintptr_t>(pts) | reinterpret_cast<uintptr_t>(FEs)) & 15) {
        set_error("%s: points and FEs must be 16-byte aligned (float4 rows)", fn);
        return PPM_ERR_INVALID_ARG;
    }
    return PPM_OK;
}

// An optional float4 output (PPs, displacements): NULL or 16-byte aligned.
static int check_out4(const char* fn, const float* out, const char* name) {
    if (reinterpret_cast<uintptr_t>(out) & 15) {
        set_error("%s: %s must be 16-byte aligned (float4 rows)", fn, name);
        return PPM_ERR_INVALID_ARG;
    }
    return PPM_OK;
}

static int check_ctl(const char* fn, const PpmRelaxCtl* c, int n_scan) {
    if (!c) { set_error("%s: ctl is NULL", fn); return PPM_ERR_INVALID_ARG; }
    if (c->max_iter < 1 || c->max_iter > (1 << 20)) {
        set_error("%s: max_iter=%d out of range [1, 2^20]", fn, c->max_iter);
        return PPM_ERR_INVALID_ARG;
    }
    if (!std::isfinite(c->f2conv) || c->f2conv < 0.f) { set_error("%s: invalid f2conv", fn); return PPM_ERR_INVALID_ARG; }
    if (c->damped != 0 && c->damped != 1) { set_error("%s: damped=%d is not 0 or 1", fn, c->damped); return PPM_ERR_INVALID_ARG; }
    if (c->lanes != 0 && c->lanes != 1 && c->lanes != 2 && c->lanes != 4) {
        set_error("%s: lanes=%d is not 0, 1, 2 or 4", fn, c->lanes);
        return PPM_ERR_INVALID_ARG;
    }
    if (c->scan_ny < 0 || (c->scan_ny > 0 && n_scan % c->scan_ny != 0)) {
        set_error("%s: scan_ny=%d does not divide n_scan=%d", fn, c->scan_ny, n_scan);
        return PPM_ERR_INVALID_ARG;
    }
    return PPM_OK;
}

}  // namespace ppm

using namespace ppm;

extern "C" {

int ppm_version(void) { return PPM_API_VERSION; }

const char* ppm_last_error(void) { return g_err; }

int ppm_stream_create(void** stream) {
    g_err[0] = 0;
    if (!stream) { set_error("ppm_stream_create: stream is NULL"); return PPM_ERR_INVALID_ARG; }
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) { set_error("ppm_stream_create: %s", hipGetErrorString(e)); return PPM_ERR_HIP; }
    *stream = s;
    return PPM_OK;
}

int ppm_stream_destroy(void* stream) {
    g_err[0] = 0;
    if (!stream) { set_error("ppm_stream_destroy: stream is NULL"); return PPM_ERR_INVALID_ARG; }
    const hipError_t e = hipStreamDestroy(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) { set_error("ppm_stream_destroy: %s", hipGetErrorString(e)); return PPM_ERR_HIP; }
    return PPM_OK;
}

static int grid_project_impl(const char* fn, int model, const float* atoms, const float* cpar, int na,
                             const PpmGridSpec* spec, const float Qs[4], const float QZs[4], int npbc_a, int npbc_b,
                             const float lvA[3], const float lvB[3], int cap, const float* base, float* grid,
                             void* stream, bool acc64 = false, int klo = 0, int nzw = 0) {
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (nzw > 0 && (klo < 0 || (long long)klo + nzw > spec->nz)) {
        set_error("%s: window [%d, %d) outside the grid's %d planes", fn, klo, klo + nzw, spec->nz);
        return PPM_ERR_INVALID_ARG;
    }
    if (model != PPM_MODEL_LJ && model != PPM_MODEL_MORSE) { set_error("%s: unknown model %d", fn, model); return PPM_ERR_INVALID_ARG; }
    if (na < 1 || na > (1 << 20)) { set_error("%s: invalid atom count %d", fn, na); return PPM_ERR_INVALID_ARG; }
    if (!atoms || !cpar || !grid || !Qs || !QZs) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    if (npbc_a < 0 || npbc_b < 0 || npbc_a > 64 || npbc_b > 64) {
        set_error("%s: image range (%d,%d) out of [0,64]", fn, npbc_a, npbc_b);
        return PPM_ERR_INVALID_ARG;
    }
    if ((npbc_a > 0 && !lvA) || (npbc_b > 0 && !lvB)) { set_error("%s: images requested without lattice vectors", fn); return PPM_ERR_INVALID_ARG; }
    if ((long long)na * (2 * npbc_a + 1) * (2 * npbc_b + 1) > (1LL << 28)) { set_error("%s: too many image sites", fn); return PPM_ERR_INVALID_ARG; }
    if (!finite4(Qs, 4) || !finite4(QZs, 4)) { set_error("%s: non-finite tip charges", fn); return PPM_ERR_INVALID_ARG; }
    return launch_grid_project(model, atoms, cpar, na, *spec, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap, base, grid,
                               acc64, (hipStream_t)stream, klo, nzw);
}

int ppm_grid_project(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec* spec,
                     const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                     const float lvB[3], int cap, float* grid, void* stream) {
    return grid_project_impl("ppm_grid_project", model, atoms, cpar, na, spec, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap,
                             nullptr, grid, stream);
}

int ppm_grid_project_window(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec* spec,
                            int klo, int nzw, const float Qs[4], const float QZs[4], int npbc_a, int npbc_b,
                            const float lvA[3], const float lvB[3], int cap, float* grid, void* stream) {
    if (nzw < 1) {
        g_err[0] = 0;
        set_error("ppm_grid_project_window: empty window (nzw=%d)", nzw);
        return PPM_ERR_INVALID_ARG;
    }
    return grid_project_impl("ppm_grid_project_window", model, atoms, cpar, na, spec, Qs, QZs, npbc_a, npbc_b, lvA,
                             lvB, cap, nullptr, grid, stream, false, klo, nzw);
}

int ppm_grid_project_acc64(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec* spec,
                           const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                           const float lvB[3], int cap, float* grid, void* stream) {
    return grid_project_impl("ppm_grid_project_acc64", model, atoms, cpar, na, spec, Qs, QZs, npbc_a, npbc_b, lvA,
                             lvB, cap, nullptr, grid, stream, true);
}

int ppm_grid_project_add(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec* spec,
                         const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                         const float lvB[3], int cap, const float* base, float* grid, void* stream) {
    const char* fn = "ppm_grid_project_add";
    if (!base || base == grid) {
        g_err[0] = 0;
        set_error("%s: base must be a separate grid buffer", fn);
        return PPM_ERR_INVALID_ARG;
    }
    return grid_project_impl(fn, model, atoms, cpar, na, spec, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap, base, grid,
                             stream);
}

int ppm_grid_project_f64(int model, const double* atoms, const double* cpar, int na, const PpmGridSpec* spec,
                         const double Qs[4], const double QZs[4], int npbc_a, int npbc_b, const double lvA[3],
                         const double lvB[3], int cap, double* grid, void* stream) {
    const char* fn = "ppm_grid_project_f64";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (model != PPM_MODEL_LJ && model != PPM_MODEL_MORSE) { set_error("%s: unknown model %d", fn, model); return PPM_ERR_INVALID_ARG; }
    if (na < 1 || na > (1 << 20)) { set_error("%s: invalid atom count %d", fn, na); return PPM_ERR_INVALID_ARG; }
    if (!atoms || !cpar || !grid || !Qs || !QZs) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    if (npbc_a < 0 || npbc_b < 0 || npbc_a > 64 || npbc_b > 64) {
        set_error("%s: image range (%d,%d) out of [0,64]", fn, npbc_a, npbc_b);
        return PPM_ERR_INVALID_ARG;
    }
    if ((npbc_a > 0 && !lvA) || (npbc_b > 0 && !lvB)) { set_error("%s: images requested without lattice vectors", fn); return PPM_ERR_INVALID_ARG; }
    if ((long long)na * (2 * npbc_a + 1) * (2 * npbc_b + 1) > (1LL << 28)) { set_error("%s: too many image sites", fn); return PPM_ERR_INVALID_ARG; }
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(Qs[k]) || !std::isfinite(QZs[k])) { set_error("%s: non-finite tip charges", fn); return PPM_ERR_INVALID_ARG; }
    if (spec->nz > 65535) { set_error("%s: nz=%d above 65535", fn, spec->nz); return PPM_ERR_INVALID_ARG; }
    return launch_grid_project_f64(model, atoms, cpar, na, *spec, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap, grid,
                                   (hipStream_t)stream);
}

int ppm_relax_strokes_tilted_f64(const double* grid, const PpmGridSpec* spec, const double* points, int n_scan,
                                 int nz, const PpmTip* tip, const PpmRelaxCtl* ctl, double* FEs, double* PPs,
                                 void* stream) {
    const char* fn = "ppm_relax_strokes_tilted_f64";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (!grid || !points || !FEs || !tip) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    if (n_scan < 1 || nz < 1 || (long long)n_scan * nz > (1LL << 31) / 4) {
        set_error("%s: invalid scan size %d x %d", fn, n_scan, nz);
        return PPM_ERR_INVALID_ARG;
    }
    if (int r = check_ctl(fn, ctl, n_scan)) return r;
    return launch_relax_tilted_f64(grid, *spec, points, n_scan, nz, *tip, *ctl, FEs, PPs, (hipStream_t)stream);
}

static int check_shape(const char* fn, const int* n) {
    if (!n) { set_error("%s: shape is NULL", fn); return PPM_ERR_INVALID_ARG; }
    for (int k = 0; k < 3; ++k)
        if (n[k] < 1 || n[k] > 16384) { set_error("%s: shape[%d]=%d out of [1,16384]", fn, k, n[k]); return PPM_ERR_INVALID_ARG; }
    if ((long long)n[0] * n[1] * n[2] > (1LL << 31) - 1) { set_error("%s: grid too large", fn); return PPM_ERR_INVALID_ARG; }
    return PPM_OK;
}

size_t ppm_fdbm_workspace_size(int nx, int ny, int nz) {
    const int n[3] = {nx, ny, nz};
    if (check_shape("ppm_fdbm_workspace_size", n)) return 0;
    return 2 * fdbm_spectrum_bytes(nx, ny, nz);
}

int ppm_fdbm_release_plans(void) {
    g_err[0] = 0;
    return release_fft_plans();
}

int ppm_fdbm_fft_combine(int mode, int nx, int ny, int nz, float step, const float* a, const float* b, float scale,
                         float* out, void* workspace, size_t workspace_bytes, void* stream) {
    const char* fn = "ppm_fdbm_fft_combine";
    g_err[0] = 0;
    const int n[3] = {nx, ny, nz};
    if (int r = check_shape(fn, n)) return r;
    if (mode < PPM_FFT_CORRELATE || mode > PPM_FFT_POISSON_CONVOLVE) { set_error("%s: unknown mode %d", fn, mode); return PPM_ERR_INVALID_ARG; }
    if (!a || !out || !workspace || (mode != PPM_FFT_POISSON && !b)) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    if (out == a || out == b) { set_error("%s: out must not alias an input", fn); return PPM_ERR_INVALID_ARG; }
    if (!(step > 0.f) || !std::isfinite(step) || !std::isfinite(scale)) { set_error("%s: invalid step or scale", fn); return PPM_ERR_INVALID_ARG; }
    if (workspace_bytes < 2 * fdbm_spectrum_bytes(nx, ny, nz)) { set_error("%s: workspace too small", fn); return PPM_ERR_INVALID_ARG; }
    return launch_fft_combine(mode, nx, ny, nz, step, a, b, scale, out, workspace, (hipStream_t)stream);
}

int ppm_fdbm_pauli_scale(float* field, long long n, float A, float beta, float clip, void* stream) {
    const char* fn = "ppm_fdbm_pauli_scale";
    g_err[0] = 0;
    if (!field || n < 1) { set_error("%s: empty or NULL field", fn); return PPM_ERR_INVALID_ARG; }
    if (!std::isfinite(A) || !std::isfinite(beta) || !(clip > 0.f)) { set_error("%s: invalid A, beta or clip", fn); return PPM_ERR_INVALID_ARG; }
    return launch_pauli_scale(field, n, A, beta, clip, (hipStream_t)stream);
}

int ppm_fdbm_pad_roll(const float* src, const int src_shape[3], float* dst, const int shape[3], const int offset[3],
                      const int roll[3], void* stream) {
    const char* fn = "ppm_fdbm_pad_roll";
    g_err[0] = 0;
    if (int r = check_shape(fn, src_shape)) return r;
    if (int r = check_shape(fn, shape)) return r;
    if (!src || !dst || !offset || !roll || src == dst) { set_error("%s: NULL or aliased pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_pad_roll(src, src_shape, dst, shape, offset, roll, (hipStream_t)stream);
}

int ppm_fdbm_flip3(const float* src, float* dst, const int shape[3], void* stream) {
    const char* fn = "ppm_fdbm_flip3";
    g_err[0] = 0;
    if (int r = check_shape(fn, shape)) return r;
    if (!src || !dst || src == dst) { set_error("%s: NULL or aliased pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_flip3(src, dst, shape, (hipStream_t)stream);
}

int ppm_fdbm_dispersion(const float* atoms, int na, const float origin[3], float step, const int shape[3], float RA,
                        int accumulate, float* out, void* stream) {
    const char* fn = "ppm_fdbm_dispersion";
    g_err[0] = 0;
    if (int r = check_shape(fn, shape)) return r;
    if (!atoms || !origin || !out || na < 1) { set_error("%s: NULL pointer or no atoms", fn); return PPM_ERR_INVALID_ARG; }
    if (!(step > 0.f) || !std::isfinite(RA) || !finite4(origin, 3)) { set_error("%s: invalid step, RA or origin", fn); return PPM_ERR_INVALID_ARG; }
    return launch_dispersion(atoms, na, origin, step, shape, RA * RA, accumulate, out, (hipStream_t)stream);
}

int ppm_fdbm_compose_gradient(const float* e1, const float* e2, const float* e3, const int shape[3], float step,
                              float* grid, void* stream) {
    const char* fn = "ppm_fdbm_compose_gradient";
    g_err[0] = 0;
    if (int r = check_shape(fn, shape)) return r;
    if (!grid || (!e1 && !e2 && !e3)) { set_error("%s: no output or no input field", fn); return PPM_ERR_INVALID_ARG; }
    if (!(step > 0.f)) { set_error("%s: invalid step", fn); return PPM_ERR_INVALID_ARG; }
    return launch_compose_gradient(e1, e2, e3, shape, step, grid, (hipStream_t)stream);
}

static int check_lattice(const char* fn, const double a[2], const double b[2]) {
    if (!a || !b) { set_error("%s: lattice vectors are NULL", fn); return PPM_ERR_INVALID_ARG; }
    for (int k = 0; k < 2; ++k)
        if (!std::isfinite(a[k]) || !std::isfinite(b[k])) { set_error("%s: non-finite lattice vector", fn); return PPM_ERR_INVALID_ARG; }
    if (!(std::fabs(a[0] * b[1] - a[1] * b[0]) > 1e-12)) { set_error("%s: degenerate lattice (|a x b| = 0)", fn); return PPM_ERR_INVALID_ARG; }
    return PPM_OK;
}

static int check_ewald_sys(const char* fn, int n_layers, int n_harm) {
    if (n_layers < 1 || n_layers > 65536) { set_error("%s: n_layers=%d out of [1,65536]", fn, n_layers); return PPM_ERR_INVALID_ARG; }
    if (n_harm < 1 || n_harm > 64) { set_error("%s: n_harm=%d out of [1,64]", fn, n_harm); return PPM_ERR_INVALID_ARG; }
    if ((long long)n_layers * ewald_num_g(n_harm) > (1LL << 31) - 1) { set_error("%s: n_layers * n_g too large", fn); return PPM_ERR_INVALID_ARG; }
    return PPM_OK;
}

int ppm_ewald_num_g(int n_harm) { return n_harm < 0 ? 0 : ewald_num_g(n_harm); }

size_t ppm_ewald_system_bytes(int n_layers, int n_harm) {
    if (check_ewald_sys("ppm_ewald_system_bytes", n_layers, n_harm)) return 0;
    return ewald_system_bytes(n_layers, n_harm);
}

size_t ppm_ewald_kept_count_offset(int n_layers, int n_harm) {
    if (check_ewald_sys("ppm_ewald_kept_count_offset", n_layers, n_harm)) return 0;
    return ewald_kept_count_offset(n_layers, n_harm);
}

int ppm_ewald_prepare(const float* ions, const int* layer_off, int n_layers, const double a[2], const double b[2],
                      int n_harm, double drop_tol, void* system, size_t system_bytes, void* stream) {
    const char* fn = "ppm_ewald_prepare";
    g_err[0] = 0;
    if (int r = check_ewald_sys(fn, n_layers, n_harm)) return r;
    if (int r = check_lattice(fn, a, b)) return r;
    if (!ions || !layer_off || !system) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    if (!(drop_tol >= 0.0) || !(drop_tol < 1.0)) { set_error("%s: drop_tol=%g out of [0,1)", fn, drop_tol); return PPM_ERR_INVALID_ARG; }
    if (system_bytes < ewald_system_bytes(n_layers, n_harm)) { set_error("%s: system buffer too small", fn); return PPM_ERR_INVALID_ARG; }
    return launch_ewald_prepare(ions, layer_off, n_layers, a, b, n_harm, drop_tol, system, (hipStream_t)stream);
}

int ppm_ewald_eval(int mode, const float* points, int np, const void* system, size_t system_bytes, int n_layers,
                   int n_harm, const double a[2], const double b[2], float* phi, void* stream) {
    const char* fn = "ppm_ewald_eval";
    g_err[0] = 0;
    if (mode != PPM_EWALD_FULL && mode != PPM_EWALD_VACUUM) { set_error("%s: unknown mode %d", fn, mode); return PPM_ERR_INVALID_ARG; }
    if (int r = check_ewald_sys(fn, n_layers, n_harm)) return r;
    if (int r = check_lattice(fn, a, b)) return r;
    if (np < 0) { set_error("%s: np=%d < 0", fn, np); return PPM_ERR_INVALID_ARG; }
    if (np > 0 && (!points || !system || !phi)) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    if (system_bytes < ewald_system_bytes(n_layers, n_harm)) { set_error("%s: system buffer too small for n_layers, n_harm", fn); return PPM_ERR_INVALID_ARG; }
    return launch_ewald_points(mode, points, np, system, n_layers, n_harm, a, b, phi, (hipStream_t)stream);
}

int ppm_ewald_brute(const float* points, int np, const float* ions, int n_ions, const double a[2], const double b[2],
                    int n_rep, int square, float* phi, void* stream) {
    const char* fn = "ppm_ewald_brute";
    g_err[0] = 0;
    if (int r = check_lattice(fn, a, b)) return r;
    if (np < 0 || n_ions < 0 || n_rep < 0 || n_rep > 4096) { set_error("%s: invalid np=%d n_ions=%d n_rep=%d", fn, np, n_ions, n_rep); return PPM_ERR_INVALID_ARG; }
    if (np > 0 && (!points || !phi || (n_ions > 0 && !ions))) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_ewald_brute(points, np, ions, n_ions, a, b, n_rep, square != 0, phi, (hipStream_t)stream);
}

int ppm_ewald_cluster(const float* points, int np, const float* ions, int n_ions, float* phi, void* stream) {
    const char* fn = "ppm_ewald_cluster";
    g_err[0] = 0;
    if (np < 0 || n_ions < 0) { set_error("%s: invalid np=%d n_ions=%d", fn, np, n_ions); return PPM_ERR_INVALID_ARG; }
    if (np > 0 && (!points || !phi || (n_ions > 0 && !ions))) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_ewald_cluster(points, np, ions, n_ions, phi, (hipStream_t)stream);
}

size_t ppm_ewald_grid_workspace_size(int n_harm, int nz) {
    if (n_harm < 1 || n_harm > 64 || nz < 1) return 0;
    return ewald_grid_workspace(n_harm, nz);
}

int ppm_ewald_tip_grid(const void* system, size_t system_bytes, int n_layers, int n_harm, const double a[2],
                       const double b[2], const PpmGridSpec* spec, const float Qs[4], const float QZs[4],
                       int accumulate, void* workspace, size_t workspace_bytes, float* grid, void* stream) {
    const char* fn = "ppm_ewald_tip_grid";
    g_err[0] = 0;
    if (int r = check_ewald_sys(fn, n_layers, n_harm)) return r;
    if (int r = check_lattice(fn, a, b)) return r;
    if (int r = check_spec(fn, spec)) return r;
    if (!system || !Qs || !QZs || !workspace || !grid) { set_error("%s: NULL pointer argument", fn); return PPM_ERR_INVALID_ARG; }
    if (!finite4(Qs, 4) || !finite4(QZs, 4)) { set_error("%s: non-finite Qs/QZs", fn); return PPM_ERR_INVALID_ARG; }
    if (spec->dA[2] != 0.f || spec->dB[2] != 0.f || spec->dC[0] != 0.f || spec->dC[1] != 0.f) {
        set_error("%s: dA, dB must lie in the xy plane and dC along z", fn);
        return PPM_ERR_INVALID_ARG;
    }
    if (spec->nz > 65535 * 8) { set_error("%s: nz=%d too large", fn, spec->nz); return PPM_ERR_INVALID_ARG; }
    if (workspace_bytes < ewald_grid_workspace(n_harm, spec->nz)) { set_error("%s: workspace too small", fn); return PPM_ERR_INVALID_ARG; }
    if (system_bytes < ewald_system_bytes(n_layers, n_harm)) { set_error("%s: system buffer too small for n_layers, n_harm", fn); return PPM_ERR_INVALID_ARG; }
    return launch_ewald_tip_grid(system, n_layers, n_harm, a, b, *spec, Qs, QZs, accumulate, workspace, grid,
                                 (hipStream_t)stream);
}

int ppm_relax_strokes_tilted(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                             const PpmTip* tip, const PpmRelaxCtl* ctl, float* FEs, float* PPs, void* stream) {
    const char* fn = "ppm_relax_strokes_tilted";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, FEs)) return r;
    if (int r = check_out4(fn, PPs, "PPs")) return r;
    if (int r = check_ctl(fn, ctl, n_scan)) return r;
    if (!grid || !tip) { set_error("%s: NULL grid/tip", fn); return PPM_ERR_INVALID_ARG; }
    return launch_relax_tilted(grid, *spec, points, n_scan, nz, *tip, *ctl, FEs, PPs, (hipStream_t)stream);
}

int ppm_relax_strokes_tilted_df(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                                const PpmTip* tip, const PpmRelaxCtl* ctl, float* FEs, float* PPs, const float* W,
                                const int* start, int taps, int nz_out, float* df, void* stream) {
    const char* fn = "ppm_relax_strokes_tilted_df";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, FEs)) return r;
    if (int r = check_out4(fn, PPs, "PPs")) return r;
    if (int r = check_ctl(fn, ctl, n_scan)) return r;
    if (!grid || !tip) { set_error("%s: NULL grid/tip", fn); return PPM_ERR_INVALID_ARG; }
    if (!W || !start || !df || taps < 1 || taps > nz || nz_out < 1) {
        set_error("%s: invalid df operator (taps=%d, nz_out=%d)", fn, taps, nz_out);
        return PPM_ERR_INVALID_ARG;
    }
    return launch_relax_tilted(grid, *spec, points, n_scan, nz, *tip, *ctl, FEs, PPs, (hipStream_t)stream, 0, 0,
                               nullptr, W, start, taps, nz_out, df);
}

int ppm_relax_strokes_tilted_window(const float* grid, const PpmGridSpec* spec, int klo, int nzw,
                                    const float* points, int n_scan, int nz, const PpmTip* tip,
                                    const PpmRelaxCtl* ctl, float* FEs, float* PPs, int* flag, void* stream) {
    const char* fn = "ppm_relax_strokes_tilted_window";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (nzw < 2 || klo < 0 || (long long)klo + nzw > spec->nz) {
        set_error("%s: window [%d, %d) outside the grid's %d planes (at least 2 planes)", fn, klo, klo + nzw, spec->nz);
        return PPM_ERR_INVALID_ARG;
    }
    if (int r = check_scan(fn, points, n_scan, nz, FEs)) return r;
    if (int r = check_out4(fn, PPs, "PPs")) return r;
    if (int r = check_ctl(fn, ctl, n_scan)) return r;
    if (!grid || !tip || !flag) { set_error("%s: NULL grid/tip/flag", fn); return PPM_ERR_INVALID_ARG; }
    return launch_relax_tilted(grid, *spec, points, n_scan, nz, *tip, *ctl, FEs, PPs, (hipStream_t)stream, klo, nzw,
                               flag);
}

int ppm_relax_strokes(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                      const float dTip[4], const float stiffness[4], const float dpos0[4], const float relax[4],
                      const PpmRelaxCtl* ctl, float* FEs, float* disps, float* PPs, void* stream) {
    const char* fn = "ppm_relax_strokes";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, FEs)) return r;
    if (int r = check_out4(fn, disps, "disps")) return r;
    if (int r = check_out4(fn, PPs, "PPs")) return r;
    if (int r = check_ctl(fn, ctl, n_scan)) return r;
    if (!grid || !dTip || !stiffness || !dpos0 || !relax) { set_error("%s: NULL argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_relax_strokes(grid, *spec, points, n_scan, nz, dTip, stiffness, dpos0, relax, *ctl, FEs, disps, PPs,
                                (hipStream_t)stream);
}

int ppm_relax_strokes_2d(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                         float k_lat, float dh, float dt, float damp, const PpmRelaxCtl* ctl, float* FEs, float* disps,
                         float* PPs, void* stream) {
    const char* fn = "ppm_relax_strokes_2d";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, FEs)) return r;
    if (int r = check_out4(fn, disps, "disps")) return r;
    if (int r = check_out4(fn, PPs, "PPs")) return r;
    if (int r = check_ctl(fn, ctl, n_scan)) return r;
    if (!grid) { set_error("%s: NULL grid", fn); return PPM_ERR_INVALID_ARG; }
    if (!std::isfinite(k_lat) || !std::isfinite(dh) || !std::isfinite(dt) || !std::isfinite(damp)) {
        set_error("%s: non-finite k_lat/dh/dt/damp", fn);
        return PPM_ERR_INVALID_ARG;
    }
    return launch_relax_strokes_2d(grid, *spec, points, n_scan, nz, k_lat, dh, dt, damp, *ctl, FEs, disps, PPs,
                                   (hipStream_t)stream);
}

int ppm_fe_in_strokes_tilted(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                             const PpmTip* tip, const float dTip[4], const float dpos0[4], float* FEs, void* stream) {
    const char* fn = "ppm_fe_in_strokes_tilted";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, FEs)) return r;
    if (!grid || !tip || !dTip || !dpos0) { set_error("%s: NULL argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_fe_in_strokes(grid, *spec, points, n_scan, nz, *tip, dTip, dpos0, FEs, (hipStream_t)stream);
}

int ppm_relax_adjoint(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                      const PpmTip* tip, const float* PPs, const float* dL_dFEs, float* grid_cot, float* lambdas,
                      void* stream) {
    const char* fn = "ppm_relax_adjoint";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, PPs)) return r;
    if (!grid || !tip || !dL_dFEs || !grid_cot || !lambdas) { set_error("%s: NULL argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_relax_adjoint(grid, *spec, points, n_scan, nz, *tip, PPs, dL_dFEs, grid_cot, lambdas,
                                (hipStream_t)stream);
}

int ppm_relax_polish(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                     const PpmTip* tip, int steps, float* PPs, float* FEs, void* stream) {
    const char* fn = "ppm_relax_polish";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, PPs)) return r;
    if (!grid || !tip) { set_error("%s: NULL grid/tip", fn); return PPM_ERR_INVALID_ARG; }
    if (steps < 0 || steps > 64) { set_error("%s: steps=%d out of [0,64]", fn, steps); return PPM_ERR_INVALID_ARG; }
    return launch_relax_polish(grid, *spec, points, n_scan, nz, *tip, steps, PPs, FEs, (hipStream_t)stream);
}

int ppm_relax_state_diag(const float* grid, const PpmGridSpec* spec, const float* points, int n_scan, int nz,
                         const PpmTip* tip, const float* PPs, float* diag, void* stream) {
    const char* fn = "ppm_relax_state_diag";
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (int r = check_scan(fn, points, n_scan, nz, PPs)) return r;
    if (!grid || !tip || !diag) { set_error("%s: NULL argument", fn); return PPM_ERR_INVALID_ARG; }
    return launch_relax_state_diag(grid, *spec, points, n_scan, nz, *tip, PPs, diag, (hipStream_t)stream);
}

static int check_direct(const char* fn, const float* atoms, const float* cMs, int na, const float* points,
                        int n_scan, int nz, const PpmTip* tip, const float* Qs, const float* QZs) {
    if (na < 1 || na > PPM_DIRECT_MAX_ATOMS) {
        set_error("%s: nAtoms=%d out of range [1,%d]", fn, na, PPM_DIRECT_MAX_ATOMS);
        return PPM_ERR_INVALID_ARG;
    }
    if (!atoms || !cMs || !points || !tip || !Qs || !QZs) { set_error("%s: NULL argument", fn); return PPM_ERR_INVALID_ARG; }
    if (n_scan < 1 || nz < 1 || (long long)n_scan * nz > (1LL << 31)) {
        set_error("%s: invalid n_scan=%d nz=%d", fn, n_scan, nz);
        return PPM_ERR_INVALID_ARG;
    }
    if (!(tip->tipC[3] < 0.f)) { set_error("%s: dtip must be <0, got %g", fn, (double)tip->tipC[3]); return PPM_ERR_INVALID_ARG; }
    if (!finite4(Qs, 4) || !finite4(QZs, 4)) { set_error("%s: non-finite tip charges", fn); return PPM_ERR_INVALID_ARG; }
    return PPM_OK;
}

int ppm_relax_direct_morse(const float* atoms, const float* cMs, int na, const float* points, int n_scan, int nz,
                           const PpmTip* tip, const float Qs[4], const float QZs[4], const PpmRelaxCtl* ctl,
                           float* FEs, float* PPs, void* stream) {
    const char* fn = "ppm_relax_direct_morse";
    g_err[0] = 0;
    if (int r = check_direct(fn, atoms, cMs, na, points, n_scan, nz, tip, Qs, QZs)) return r;
    if (int r = check_ctl(fn, ctl, n_scan)) return r;
    if (!FEs) { set_error("%s: NULL FEs", fn); return PPM_ERR_INVALID_ARG; }
    return launch_direct_relax(atoms, cMs, na, points, n_scan, nz, *tip, Qs, QZs, *ctl, FEs, PPs, (hipStream_t)stream);
}

size_t ppm_direct_morse_vjp_workspace_size(int na, int n_scan, int nz) {
    if (na < 1 || na > PPM_DIRECT_MAX_ATOMS || n_scan < 1 || nz < 1) return 0;
    return direct_vjp_workspace(na, n_scan, nz);
}

int ppm_direct_morse_vjp(const float* atoms, const float* cMs, int na, const float* points, int n_scan, int nz,
                         const PpmTip* tip, const float Qs[4], const float QZs[4], const float* PPs,
                         const float* dL_dFEs, float* lambdas, float* diag, void* workspace, size_t workspace_bytes,
                         double* grad, void* stream) {
    const char* fn = "ppm_direct_morse_vjp";
    g_err[0] = 0;
    if (int r = check_direct(fn, atoms, cMs, na, points, n_scan, nz, tip, Qs, QZs)) return r;
    if (!PPs || !dL_dFEs || !lambdas || !diag || !workspace || !grad) {
        set_error("%s: NULL argument", fn);
        return PPM_ERR_INVALID_ARG;
    }
    const size_t need = direct_vjp_workspace(na, n_scan, nz);
    if (workspace_bytes < need) {
        set_error("%s: workspace of %zu bytes < required %zu (ppm_direct_morse_vjp_workspace_size)", fn,
                  workspace_bytes, need);
        return PPM_ERR_INVALID_ARG;
    }
    return launch_direct_vjp(atoms, cMs, na, points, n_scan, nz, *tip, Qs, QZs, PPs, dL_dFEs, lambdas, diag, workspace,
                             grad, (hipStream_t)stream);
}

size_t ppm_batch_workspace_size(int nmol) { return nmol < 1 ? 0 : batch_workspace(nmol); }

static int check_batch(const char* fn, const PpmMolDesc* mols, const PpmBatchParams* params, const float* atoms,
                       const float* cpar, const float* grids, const float* FEs, const float* df, const float* dfW,
                       const int* dfStart, int workers, int reset) {
    if (!mols || !params || !atoms || !cpar || !grids || !FEs) {
        set_error("%s: NULL argument", fn);
        return PPM_ERR_INVALID_ARG;
    }
    const PpmBatchParams& P = *params;
    if (P.model != PPM_MODEL_LJ && P.model != PPM_MODEL_MORSE) { set_error("%s: unknown model %d", fn, P.model); return PPM_ERR_INVALID_ARG; }
    if (P.nmol < 1 || P.nitems < 1 || P.nz < 1) { set_error("%s: empty batch", fn); return PPM_ERR_INVALID_ARG; }
    if (P.vz != 4 && P.vz != 8) { set_error("%s: vz=%d not in {4,8}", fn, P.vz); return PPM_ERR_INVALID_ARG; }
    if (P.tile_w != 4 && P.tile_w != 8 && P.tile_w != 16) { set_error("%s: tile_w=%d not in {4,8,16}", fn, P.tile_w); return PPM_ERR_INVALID_ARG; }
    if (P.mode < PPM_BATCH_FUSED || P.mode > PPM_BATCH_RELAX_ONLY) { set_error("%s: unknown mode %d", fn, P.mode); return PPM_ERR_INVALID_ARG; }
    if (int r = check_ctl(fn, &P.ctl, 1)) return r;
    if (df && (!dfW || !dfStart || P.df_taps < 1 || P.df_taps > P.nz || P.df_nz_out < 1)) {
        set_error("%s: df requested with an invalid operator", fn);
        return PPM_ERR_INVALID_ARG;
    }
    if (workers < 1 || workers > (1 << 16)) { set_error("%s: workers=%d out of range", fn, workers); return PPM_ERR_INVALID_ARG; }
    if (reset < 0 || reset > 2) { set_error("%s: reset=%d not in {0,1,2}", fn, reset); return PPM_ERR_INVALID_ARG; }
    return PPM_OK;
}

int ppm_batch_forward(const PpmMolDesc* mols, const PpmBatchParams* params, const float* atoms, const float* cpar,
                      float* grids, float* FEs, float* PPs, float* df, const float* dfW, const int* dfStart,
                      void* workspace, size_t workspace_bytes, int workers, int reset, void* stream) {
    const char* fn = "ppm_batch_forward";
    g_err[0] = 0;
    if (int r = check_batch(fn, mols, params, atoms, cpar, grids, FEs, df, dfW, dfStart, workers, reset)) return r;
    if (!workspace) { set_error("%s: NULL workspace", fn); return PPM_ERR_INVALID_ARG; }
    if (workspace_bytes < batch_workspace(params->nmol)) { set_error("%s: workspace too small", fn); return PPM_ERR_INVALID_ARG; }
    return launch_batch(mols, *params, atoms, cpar, grids, FEs, PPs, df, dfW, dfStart, workspace, workers, reset,
                        (hipStream_t)stream);
}

int ppm_batch_forward_noalloc(const PpmMolDesc* mols, const PpmBatchParams* params, const float* atoms,
                              const float* cpar, float* grids, float* FEs, float* PPs, float* df, const float* dfW,
                              const int* dfStart, int* status, int workers, int reset, void* stream) {
    const char* fn = "ppm_batch_forward_noalloc";
    g_err[0] = 0;
    if (int r = check_batch(fn, mols, params, atoms, cpar, grids, FEs, df, dfW, dfStart, workers, reset)) return r;
    if (!status) { set_error("%s: NULL status", fn); return PPM_ERR_INVALID_ARG; }
    if (reset == 0) {
        set_error("%s: reset=0 would reuse the previous call's queue; use 1 (or 2 after a GRID_ONLY launch)", fn);
        return PPM_ERR_INVALID_ARG;
    }
    if (params->nmol > kBatchNoallocMaxMol) {
        set_error("%s: nmol=%d exceeds the allocation-free maximum of %d molecules (use ppm_batch_forward)", fn,
                  params->nmol, kBatchNoallocMaxMol);
        return PPM_ERR_INVALID_ARG;
    }
    return launch_batch_noalloc(mols, *params, atoms, cpar, grids, FEs, PPs, df, dfW, dfStart, status, workers, reset,
                                (hipStream_t)stream);
}

size_t ppm_grid_adjoint_workspace_size(int na, const PpmGridSpec* spec) {
    if (!spec || na < 1 || check_spec("ppm_grid_adjoint_workspace_size", spec) != PPM_OK) return 0;
    return grid_adjoint_workspace(na, *spec);
}

static int grid_adjoint_impl(const char* fn, int model, const float* atoms, const float* cpar, int na,
                             const PpmGridSpec* spec, const float Qs[4], const float QZs[4], int npbc_a, int npbc_b,
                             const float lvA[3], const float lvB[3], int cap, const float* grid,
                             const float* grid_cot, const float* base, void* workspace, size_t workspace_bytes,
                             int params, double* grad, void* stream) {
    g_err[0] = 0;
    if (int r = check_spec(fn, spec)) return r;
    if (model != PPM_MODEL_LJ && model != PPM_MODEL_MORSE) { set_error("%s: unknown model %d", fn, model); return PPM_ERR_INVALID_ARG; }
    if (na < 1 || na > (1 << 20)) { set_error("%s: invalid atom count %d", fn, na); return PPM_ERR_INVALID_ARG; }
    if (!atoms || !cpar || !grid || !grid_cot || !grad || !workspace || !Qs || !QZs) {
        set_error("%s: NULL pointer argument", fn);
        return PPM_ERR_INVALID_ARG;
    }
    if (npbc_a < 0 || npbc_b < 0 || npbc_a > 64 || npbc_b > 64) { set_error("%s: image range out of [0,64]", fn); return PPM_ERR_INVALID_ARG; }
    if ((npbc_a > 0 && !lvA) || (npbc_b > 0 && !lvB)) { set_error("%s: images requested without lattice vectors", fn); return PPM_ERR_INVALID_ARG; }
    if (params != 0 && params != 1) { set_error("%s: params must be 0 or 1, got %d", fn, params); return PPM_ERR_INVALID_ARG; }
    const size_t need = grid_adjoint_workspace(na, *spec);
    if (workspace_bytes < need) {
        set_error("%s: workspace of %zu bytes < required %zu (ppm_grid_adjoint_workspace_size)", fn, workspace_bytes, need);
        return PPM_ERR_INVALID_ARG;
    }
    return launch_grid_adjoint(model, atoms, cpar, na, *spec, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap, grid, grid_cot,
                               base, workspace, params, grad, (hipStream_t)stream);
}

int ppm_grid_adjoint(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec* spec,
                     const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                     const float lvB[3], int cap, const float* grid, const float* grid_cot, void* workspace,
                     size_t workspace_bytes, int params, double* grad, void* stream) {
    return grid_adjoint_impl("ppm_grid_adjoint", model, atoms, cpar, na, spec, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap,
                             grid, grid_cot, nullptr, workspace, workspace_bytes, params, grad, stream);
}

int ppm_grid_adjoint_add(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec* spec,
                         const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                         const float lvB[3], int cap, const float* base, const float* grid, const float* grid_cot,
                         void* workspace, size_t workspace_bytes, int params, double* grad, void* stream) {
    const char* fn = "ppm_grid_adjoint_add";
    if (!base) {
        g_err[0] = 0;
        set_error("%s: NULL base", fn);
        return PPM_ERR_INVALID_ARG;
    }
    return grid_adjoint_impl(fn, model, atoms, cpar, na, spec, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap, grid, grid_cot,
                             base, workspace, workspace_bytes, params, grad, stream);
}

int ppm_z_stencil(const float* in, int in_stride, int in_offset, int n_pix, int nz_in, const float* W,
                  const int* start, int taps, int nz_out, float* out, void* stream) {
    const char* fn = "ppm_z_stencil";
    g_err[0] = 0;
    if (!in || !W || !start || !out) { set_error("%s: NULL argument", fn); return PPM_ERR_INVALID_ARG; }
    if (n_pix < 1 || nz_in < 1 || nz_out < 1 || taps < 1 || taps > nz_in) {
        set_error("%s: invalid sizes n_pix=%d nz_in=%d nz_out=%d taps=%d", fn, n_pix, nz_in, nz_out, taps);
        return PPM_ERR_INVALID_ARG;
    }
    if (in_stride < 1 || in_offset < 0 || in_offset >= in_stride) {
        set_error("%s: invalid stride/offset (%d,%d)", fn, in_stride, in_offset);
        return PPM_ERR_INVALID_ARG;
    }
    const float chw[4] = {1.f, 0.f, 0.f, 0.f};
    return launch_axis_operator(in + in_offset, (long long)nz_in * in_stride, in_stride, 0, chw, 1, n_pix, nz_in, 1,
                                W, start, taps, nz_out, 1.0f, 0, out, (hipStream_t)stream);
}

int ppm_axis_operator(const float* in, long long in_outer_stride, long long in_axis_stride,
                      long long in_inner_stride, const float chw[4], int nch, int n_outer, int n_axis_in,
                      int n_inner, const float* W, const int* start, int taps, int n_axis_out, float alpha,
                      int accumulate, float* out, void* stream) {
    const char* fn = "ppm_axis_operator";
    g_err[0] = 0;
    if (!in || !W || !start || !out || !chw) { set_error("%s: NULL argument", fn); return PPM_ERR_INVALID_ARG; }
    if (n_outer < 1 || n_axis_in < 1 || n_inner < 1 || n_axis_out < 1 || taps < 1 || taps > n_axis_in) {
        set_error("%s: invalid sizes outer=%d axis_in=%d inner=%d axis_out=%d taps=%d", fn, n_outer, n_axis_in,
                  n_inner, n_axis_out, taps);
        return PPM_ERR_INVALID_ARG;
    }
    if (nch < 1 || nch > 4 || in_outer_stride < 0 || in_axis_stride < 0 || in_inner_stride < 0) {
        set_error("%s: invalid channel count or strides", fn);
        return PPM_ERR_INVALID_ARG;
    }
    if ((long long)n_outer * n_axis_out * n_inner > (1LL << 40)) { set_error("%s: output too large", fn); return PPM_ERR_INVALID_ARG; }
    return launch_axis_operator(in, in_outer_stride, in_axis_stride, in_inner_stride, chw, nch, n_outer, n_axis_in,
                                n_inner, W, start, taps, n_axis_out, alpha, accumulate, out, (hipStream_t)stream);
}

}  // extern "C"
