// ppm_fdbm.hip — FDBM force-field producer: FFT Pauli overlap, FFT Poisson / electrostatic
// convolution, London dispersion, energy composition and the central-difference force grid.
//
// Follows the reference's GPU FDBM stage 3 (spammm/SPM/AFM.py:3304-3566 _FDBMGpyFFT,
// :4204-4253 stage3_fdbm_fields_fast; kernels/AFM.cl:1151-1303 gradient_central_diff and the
// fdbm_* kernels, :1066-1113 evalDispersion_toImg with getLondon :804-814):
//   overlap(R) = dV * IFFT( FFT(rho) . conj(FFT(tip_tot)) )              (tip rolled to index 0)
//   E_pauli    = A * max(overlap, 1e-30)^beta
//   E_es       = dV * IFFT( FFT(rho_diff) . FFT(flip(tip_del)) . 4 pi C / k^2 ),  k = 0 -> 0
//   E_vdw      = - sum_a C6eff_a / (|r - r_a|^2 + RA^2)^3
//   F          = -(E[i+1] - E[i-1]) / (2 step) per axis, periodic; grid = (F, E)
// Scalar fields are (nx,ny,nz) C-order (z fastest), the reference's host layout; the force
// grid is written in the (nz,ny,nx,4) x-fastest layout the relaxation kernels read.
//
// FFTs are real-to-complex / complex-to-real hipFFT plans (FP32, as the reference's
// complex64), cached per shape and device; the half spectrum (nx, ny, nz/2+1) is combined in
// one pointwise kernel (k^2 from the index, no k-grid in memory).  Everything else is
// memory-bound elementwise / stencil work: one thread per voxel, z-fastest coalescing.
#include "ppm_common.hpp"

#include <hipfft/hipfft.h>

#include <mutex>
#include <vector>

namespace ppm {
namespace {

__device__ __forceinline__ int freq(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }

// a[i] <- a[i] op b[i] (x k-factor) x scale, over the half spectrum (nx, ny, nzh).
__global__ __launch_bounds__(256) void k_spectrum(float2* __restrict__ a, const float2* __restrict__ b, int mode,
                                                  int nx, int ny, int nz, int nzh, float kx0, float ky0, float kz0,
                                                  float scale) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)nx * ny * nzh;
    if (i >= n) return;
    const int iz = (int)(i % nzh);
    const int iy = (int)((i / nzh) % ny);
    const int ix = (int)(i / ((long long)nzh * ny));
    float2 v = a[i];
    if (mode == PPM_FFT_CORRELATE || mode == PPM_FFT_CONVOLVE || mode == PPM_FFT_POISSON_CONVOLVE) {
        const float2 w = b[i];
        const float wy = mode == PPM_FFT_CORRELATE ? -w.y : w.y;
        v = make_float2(v.x * w.x - v.y * wy, v.x * wy + v.y * w.x);
    }
    if (mode == PPM_FFT_POISSON || mode == PPM_FFT_POISSON_CONVOLVE) {
        const float kx = kx0 * (float)freq(ix, nx), ky = ky0 * (float)freq(iy, ny), kz = kz0 * (float)iz;
        const float k2 = kx * kx + ky * ky + kz * kz;
        const float s = i == 0 ? 0.f : 1.0f / k2;   // k = 0 (neutralising background) -> 0
        v.x *= s;
        v.y *= s;
    }
    a[i] = make_float2(v.x * scale, v.y * scale);
}

__global__ __launch_bounds__(256) void k_pauli_scale(float* __restrict__ f, long long n, float A, float beta,
                                                     float clip) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float o = fmaxf(f[i], clip);
    f[i] = A * powf(o, beta);
}

// dst[i] = padded[(i + r) mod n]; padded = zeros except the source block at (ox, oy, oz).
__global__ __launch_bounds__(256) void k_pad_roll(const float* __restrict__ src, int sx, int sy, int sz,
                                                  float* __restrict__ dst, int nx, int ny, int nz, int ox, int oy,
                                                  int oz, int rx, int ry, int rz) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nx * ny * nz) return;
    const int iz = (int)(i % nz), iy = (int)((i / nz) % ny), ix = (int)(i / ((long long)nz * ny));
    const int jx = ((ix + rx) % nx + nx) % nx, jy = ((iy + ry) % ny + ny) % ny, jz = ((iz + rz) % nz + nz) % nz;
    const int lx = jx - ox, ly = jy - oy, lz = jz - oz;
    float v = 0.f;
    if (lx >= 0 && lx < sx && ly >= 0 && ly < sy && lz >= 0 && lz < sz) v = src[((size_t)lx * sy + ly) * sz + lz];
    dst[i] = v;
}

// dst[ix,iy,iz] = src[nx-1-ix, ny-1-iy, nz-1-iz]
__global__ __launch_bounds__(256) void k_flip3(const float* __restrict__ src, float* __restrict__ dst, int nx, int ny,
                                               int nz) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)nx * ny * nz;
    if (i >= n) return;
    dst[i] = src[n - 1 - i];
}

// out[ix,iy,iz] (+)= - sum_a C6eff_a / (|p - r_a|^2 + RA^2)^3,  p = origin + step*(ix,iy,iz)
__global__ __launch_bounds__(256) void k_dispersion(const float4* __restrict__ atoms, int na, float ox, float oy,
                                                    float oz, float step, int nx, int ny, int nz, float ra2,
                                                    int accumulate, float* __restrict__ out) {
    __shared__ float4 sa[256];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)nx * ny * nz;
    const long long ic = i < n ? i : n - 1;
    const int iz = (int)(ic % nz), iy = (int)((ic / nz) % ny), ix = (int)(ic / ((long long)nz * ny));
    const float px = ox + step * (float)ix, py = oy + step * (float)iy, pz = oz + step * (float)iz;
    float e = 0.f;
    for (int t0 = 0; t0 < na; t0 += 256) {
        __syncthreads();
        if (t0 + (int)threadIdx.x < na) sa[threadIdx.x] = atoms[t0 + threadIdx.x];
        __syncthreads();
        const int cnt = min(256, na - t0);
        for (int j = 0; j < cnt; ++j) {
            const float4 a = sa[j];
            const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
            const float ir2 = 1.0f / (dx * dx + dy * dy + dz * dz + ra2);
            e -= a.w * (ir2 * ir2 * ir2);
        }
    }
    if (i < n) out[i] = accumulate ? out[i] + e : e;
}

// F = -(E[+1] - E[-1]) / (2 step) per axis with periodic neighbours, E = e1 + e2 + e3 (any may be
// NULL); written as float4 (Fx,Fy,Fz,E) in the (nz,ny,nx,4) layout.
__global__ __launch_bounds__(256) void k_compose_gradient(const float* __restrict__ e1, const float* __restrict__ e2,
                                                          const float* __restrict__ e3, int nx, int ny, int nz,
                                                          float step, float4* __restrict__ grid) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nx * ny * nz) return;
    const int iz = (int)(i % nz), iy = (int)((i / nz) % ny), ix = (int)(i / ((long long)nz * ny));
    auto E = [&](int x, int y, int z) {
        const size_t k = ((size_t)x * ny + y) * nz + z;
        float s = 0.f;
        if (e1) s += e1[k];
        if (e2) s += e2[k];
        if (e3) s += e3[k];
        return s;
    };
    const int xm = ix == 0 ? nx - 1 : ix - 1, xp = ix == nx - 1 ? 0 : ix + 1;
    const int ym = iy == 0 ? ny - 1 : iy - 1, yp = iy == ny - 1 ? 0 : iy + 1;
    const int zm = iz == 0 ? nz - 1 : iz - 1, zp = iz == nz - 1 ? 0 : iz + 1;
    const float h2 = 2.0f * step;
    const float fx = -(E(xp, iy, iz) - E(xm, iy, iz)) / h2;
    const float fy = -(E(ix, yp, iz) - E(ix, ym, iz)) / h2;
    const float fz = -(E(ix, iy, zp) - E(ix, iy, zm)) / h2;
    grid[((size_t)iz * ny + iy) * nx + ix] = make_float4(fx, fy, fz, E(ix, iy, iz));
}

// hipFFT plans per (device, shape), created on first use and kept until ppm_fdbm_release_plans.
// A plan's stream is state of the plan, so one caller's hipfftSetStream and its hipfftExec* calls
// run under g_plan_mu as a unit (calls on other streams wait for the enqueue, not for the GPU).
struct Plans {
    int dev, nx, ny, nz;
    hipfftHandle r2c, c2r;
};
std::mutex g_plan_mu;
std::vector<Plans> g_plans;

// The caller holds g_plan_mu.
int get_plans(int nx, int ny, int nz, hipStream_t stream, hipfftHandle* r2c, hipfftHandle* c2r) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { set_error("fdbm: hipGetDevice failed"); return PPM_ERR_HIP; }
    for (const Plans& p : g_plans)
        if (p.dev == dev && p.nx == nx && p.ny == ny && p.nz == nz) {
            *r2c = p.r2c;
            *c2r = p.c2r;
            return hipfftSetStream(*r2c, stream) == HIPFFT_SUCCESS && hipfftSetStream(*c2r, stream) == HIPFFT_SUCCESS
                       ? PPM_OK : (set_error("fdbm: hipfftSetStream failed"), PPM_ERR_HIP);
        }
    Plans p{dev, nx, ny, nz, 0, 0};
    if (hipfftPlan3d(&p.r2c, nx, ny, nz, HIPFFT_R2C) != HIPFFT_SUCCESS) {
        set_error("fdbm: hipfftPlan3d R2C (%d,%d,%d) failed", nx, ny, nz);
        return PPM_ERR_HIP;
    }
    if (hipfftPlan3d(&p.c2r, nx, ny, nz, HIPFFT_C2R) != HIPFFT_SUCCESS) {
        hipfftDestroy(p.r2c);
        set_error("fdbm: hipfftPlan3d C2R (%d,%d,%d) failed", nx, ny, nz);
        return PPM_ERR_HIP;
    }
    g_plans.push_back(p);
    *r2c = p.r2c;
    *c2r = p.c2r;
    if (hipfftSetStream(*r2c, stream) != HIPFFT_SUCCESS || hipfftSetStream(*c2r, stream) != HIPFFT_SUCCESS) {
        set_error("fdbm: hipfftSetStream failed");
        return PPM_ERR_HIP;
    }
    return PPM_OK;
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

size_t fdbm_spectrum_bytes(int nx, int ny, int nz) { return (size_t)nx * ny * (nz / 2 + 1) * sizeof(float2); }

int release_fft_plans() {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    int cur = 0;
    const bool have_cur = hipGetDevice(&cur) == hipSuccess;
    int rc = PPM_OK;
    for (const Plans& p : g_plans) {
        if (hipSetDevice(p.dev) != hipSuccess) { rc = PPM_ERR_HIP; continue; }
        if (hipfftDestroy(p.r2c) != HIPFFT_SUCCESS || hipfftDestroy(p.c2r) != HIPFFT_SUCCESS) rc = PPM_ERR_HIP;
    }
    g_plans.clear();
    if (have_cur) (void)hipSetDevice(cur);
    if (rc) set_error("fdbm: destroying a cached hipFFT plan failed");
    return rc;
}

int launch_fft_combine(int mode, int nx, int ny, int nz, float step, const float* a, const float* b, float scale,
                       float* out, void* workspace, hipStream_t stream) {
    std::lock_guard<std::mutex> lock(g_plan_mu);
    hipfftHandle r2c, c2r;
    if (int r = get_plans(nx, ny, nz, stream, &r2c, &c2r)) return r;
    const size_t half = fdbm_spectrum_bytes(nx, ny, nz);
    float2* A = reinterpret_cast<float2*>(workspace);
    float2* B = reinterpret_cast<float2*>(reinterpret_cast<char*>(workspace) + half);
    if (hipfftExecR2C(r2c, const_cast<float*>(a), reinterpret_cast<hipfftComplex*>(A)) != HIPFFT_SUCCESS) {
        set_error("fdbm: forward FFT failed");
        return PPM_ERR_HIP;
    }
    if (mode != PPM_FFT_POISSON) {
        if (hipfftExecR2C(r2c, const_cast<float*>(b), reinterpret_cast<hipfftComplex*>(B)) != HIPFFT_SUCCESS) {
            set_error("fdbm: forward FFT failed");
            return PPM_ERR_HIP;
        }
    }
    const int nzh = nz / 2 + 1;
    const double two_pi = 6.283185307179586;
    const long long nh = (long long)nx * ny * nzh;
    // the inverse transform is unnormalised: fold 1/(nx ny nz) into the spectrum scale
    const float sc = (float)((double)scale / ((double)nx * ny * nz));
    hipLaunchKernelGGL(k_spectrum, dim3(blocks_for(nh)), dim3(256), 0, stream, A, B, mode, nx, ny, nz, nzh,
                       (float)(two_pi / (nx * (double)step)), (float)(two_pi / (ny * (double)step)),
                       (float)(two_pi / (nz * (double)step)), sc);
    if (int r = check_launch("fdbm_spectrum")) return r;
    if (hipfftExecC2R(c2r, reinterpret_cast<hipfftComplex*>(A), out) != HIPFFT_SUCCESS) {
        set_error("fdbm: inverse FFT failed");
        return PPM_ERR_HIP;
    }
    return PPM_OK;
}

int launch_pauli_scale(float* field, long long n, float A, float beta, float clip, hipStream_t stream) {
    hipLaunchKernelGGL(k_pauli_scale, dim3(blocks_for(n)), dim3(256), 0, stream, field, n, A, beta, clip);
    return check_launch("fdbm_pauli_scale");
}

int launch_pad_roll(const float* src, const int s[3], float* dst, const int n[3], const int o[3], const int r[3],
                    hipStream_t stream) {
    hipLaunchKernelGGL(k_pad_roll, dim3(blocks_for((long long)n[0] * n[1] * n[2])), dim3(256), 0, stream, src, s[0],
                       s[1], s[2], dst, n[0], n[1], n[2], o[0], o[1], o[2], r[0], r[1], r[2]);
    return check_launch("fdbm_pad_roll");
}

int launch_flip3(const float* src, float* dst, const int n[3], hipStream_t stream) {
    hipLaunchKernelGGL(k_flip3, dim3(blocks_for((long long)n[0] * n[1] * n[2])), dim3(256), 0, stream, src, dst, n[0],
                       n[1], n[2]);
    return check_launch("fdbm_flip3");
}

int launch_dispersion(const float* atoms, int na, const float origin[3], float step, const int n[3], float ra2,
                      int accumulate, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_dispersion, dim3(blocks_for((long long)n[0] * n[1] * n[2])), dim3(256), 0, stream,
                       reinterpret_cast<const float4*>(atoms), na, origin[0], origin[1], origin[2], step, n[0], n[1],
                       n[2], ra2, accumulate, out);
    return check_launch("fdbm_dispersion");
}

int launch_compose_gradient(const float* e1, const float* e2, const float* e3, const int n[3], float step,
                            float* grid, hipStream_t stream) {
    hipLaunchKernelGGL(k_compose_gradient, dim3(blocks_for((long long)n[0] * n[1] * n[2])), dim3(256), 0, stream, e1,
                       e2, e3, n[0], n[1], n[2], step, reinterpret_cast<float4*>(grid));
    return check_launch("fdbm_compose_gradient");
}

}  // namespace ppm
