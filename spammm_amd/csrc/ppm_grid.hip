// ppm_grid.hip — force-field grid projection for gfx950 (stage a).
//
// Computes, for every voxel of a (nx,ny,nz) grid, the probe's sample force and
// energy (Fx,Fy,Fz,E) from all atoms and their periodic x,y images:
//   LJ     getLJ         kernels/AFM.cl:738-745
//   Morse  getMorse      kernels/AFM.cl:747-757
//   charge getCoulombAFM kernels/AFM.cl:730-736, one term per tip charge site
//          (Qs[k], z offset QZs[k]); COULOMB_CONST applied as in kernels/AFM.cl:961
//   cap    |E|>100 => all four channels scaled by 100/|E|   kernels/AFM.cl:984-985
//   images atom shifted by ix*lvA + iy*lvB, ix,iy in [-npbc,npbc] (kernels/gridFF.cl:960-996)
//
// Design (FP32 VALU-bound):
//   * one workgroup = 256 columns (ia,ib) x V consecutive z voxels per thread;
//     voxels of one thread share dx, dy and dx^2+dy^2 for every site;
//   * the atom x image "sites" are expanded once per 256-site tile into LDS by the
//     workgroup (one site per thread), then read as broadcasts by every lane;
//   * the tip charge at z offset 0 shares r^2 with the pair law (one rsqrt for both);
//     charge sites with a nonzero offset are evaluated separately;
//   * stores are float4 per voxel, consecutive lanes -> consecutive x (1 KiB per wave).
#include "ppm_common.hpp"

namespace ppm {
namespace {

constexpr int kTile = 256;
constexpr float kLog2e = 1.4426950408889634f;

struct __align__(16) Site {
    float x, y, z, q;
    float c0, c1, c2, c3;  // LJ: c6, c12, 6*c6, 12*c12   Morse: -K*log2e*R0, E0, K*log2e, -2K
};

struct GridArgs {
    int nx, ny, nz, ncol;
    float p0[3], dA[3], dB[3], dC[3];
    int na, nimg_a, nimg_b, nsites;
    int npbc_a, npbc_b;
    float lvA[3], lvB[3];
    float q0;          // COULOMB * Qs of the charge site at z offset 0 (0: none)
    int noff;          // number of charge sites with nonzero z offset (0..4)
    float qoff[4];     // COULOMB * Qs of those sites
    float zoff[4];     // their z offsets
    int cap;
};

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }

template <int MODEL>
__device__ __forceinline__ void load_site(const float4* __restrict__ atoms, const float* __restrict__ cpar,
                                          const GridArgs& a, int s, Site& out) {
    const int atom = s / (a.nimg_a * a.nimg_b);
    const int img = s - atom * (a.nimg_a * a.nimg_b);
    const float ix = (float)(img % a.nimg_a - a.npbc_a);
    const float iy = (float)(img / a.nimg_a - a.npbc_b);
    const float4 at = atoms[atom];
    out.x = at.x + ix * a.lvA[0] + iy * a.lvB[0];
    out.y = at.y + ix * a.lvA[1] + iy * a.lvB[1];
    out.z = at.z + ix * a.lvA[2] + iy * a.lvB[2];
    out.q = at.w;
    if (MODEL == PPM_MODEL_LJ) {
        const float c6 = cpar[2 * atom], c12 = cpar[2 * atom + 1];
        out.c0 = c6; out.c1 = c12; out.c2 = 6.0f * c6; out.c3 = 12.0f * c12;
    } else {
        const float R0 = cpar[4 * atom], E0 = cpar[4 * atom + 1], K = cpar[4 * atom + 2];
        const float Kl = K * kLog2e;
        out.c0 = -Kl * R0; out.c1 = E0; out.c2 = Kl; out.c3 = -2.0f * K;
    }
}

// V voxels per thread along c (z); DCZ: dC has no x,y component (every grid set-up of
// spammm/SPM/AFM.py:772-893 has dC = (0,0,L/nz)), so all V voxels share x,y.
template <int MODEL, bool HASQ0, int V, bool DCZ>
__global__ __launch_bounds__(256) void k_grid_project(const float4* __restrict__ atoms,
                                                      const float* __restrict__ cpar, GridArgs a,
                                                      float4* __restrict__ out) {
    __shared__ Site sites[kTile];
    const int col = blockIdx.x * 256 + threadIdx.x;
    const bool colok = col < a.ncol;
    const int cc = colok ? col : a.ncol - 1;
    const int ia = cc % a.nx, ib = cc / a.nx;
    const int ic0 = blockIdx.y * V;

    float px[DCZ ? 1 : V], py[DCZ ? 1 : V], pz[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const float fic = (float)(ic0 + k), fia = (float)ia, fib = (float)ib;
        pz[k] = a.p0[2] + a.dA[2] * fia + a.dB[2] * fib + a.dC[2] * fic;
        if (!DCZ || k == 0) {
            px[k] = a.p0[0] + a.dA[0] * fia + a.dB[0] * fib + a.dC[0] * fic;
            py[k] = a.p0[1] + a.dA[1] * fia + a.dB[1] * fib + a.dC[1] * fic;
        }
    }
    float fx[V], fy[V], fz[V], fe[V];
#pragma unroll
    for (int k = 0; k < V; ++k) { fx[k] = 0.f; fy[k] = 0.f; fz[k] = 0.f; fe[k] = 0.f; }

    for (int base = 0; base < a.nsites; base += kTile) {
        __syncthreads();
        const int s = base + threadIdx.x;
        if (s < a.nsites) load_site<MODEL>(atoms, cpar, a, s, sites[threadIdx.x]);
        __syncthreads();
        const int cnt = min(kTile, a.nsites - base);
        for (int j = 0; j < cnt; ++j) {
            const Site S = sites[j];
            const float qq0 = S.q * a.q0;
            float dxs[DCZ ? 1 : V], dys[DCZ ? 1 : V], dxy2[DCZ ? 1 : V];
#pragma unroll
            for (int k = 0; k < (DCZ ? 1 : V); ++k) {
                dxs[k] = px[k] - S.x;
                dys[k] = py[k] - S.y;
                dxy2[k] = fmaf(dxs[k], dxs[k], fmaf(dys[k], dys[k], kR2SAFE));
            }
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int kk = DCZ ? 0 : k;
                const float dx = dxs[kk], dy = dys[kk];
                const float dz = pz[k] - S.z;
                const float r2 = fmaf(dz, dz, dxy2[kk]);
                const float ir = rsq(r2);
                const float ir2 = ir * ir;
                float fr, e;
                if (MODEL == PPM_MODEL_LJ) {
                    const float ir6 = ir2 * ir2 * ir2;
                    e = fmaf(S.c1, ir6, -S.c0) * ir6;
                    fr = fmaf(S.c3, ir6, -S.c2) * ir6 * ir2;
                } else {
                    const float r = r2 * ir;
                    const float ex = ex2(fmaf(S.c2, r, S.c0));
                    const float Ae = S.c1 * ex;
                    e = (ex - 2.0f) * Ae;
                    fr = S.c3 * (Ae * (ex - 1.0f)) * ir;
                }
                if (HASQ0) {
                    const float ec = qq0 * ir;
                    fr = fmaf(ec, ir2, fr);
                    e += ec;
                }
                fx[k] = fmaf(fr, dx, fx[k]);
                fy[k] = fmaf(fr, dy, fy[k]);
                fz[k] = fmaf(fr, dz, fz[k]);
                fe[k] += e;
            }
            // tip charge sites with a nonzero z offset (none for a single-charge probe)
            for (int o = 0; o < a.noff; ++o) {
                const float qo = S.q * a.qoff[o];
                const float zo = a.zoff[o];
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const int kk = DCZ ? 0 : k;
                    const float dzo = (pz[k] - S.z) + zo;
                    const float iro = rsq(fmaf(dzo, dzo, dxy2[kk]));
                    const float eo = qo * iro;
                    const float fro = eo * iro * iro;
                    fx[k] = fmaf(fro, dxs[kk], fx[k]);
                    fy[k] = fmaf(fro, dys[kk], fy[k]);
                    fz[k] = fmaf(fro, dzo, fz[k]);
                    fe[k] += eo;
                }
            }
        }
    }
    if (!colok) return;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int ic = ic0 + k;
        if (ic >= a.nz) break;
        float4 v = make_float4(fx[k], fy[k], fz[k], fe[k]);
        if (a.cap) {
            const float renorm = 100.0f / fabsf(v.w);
            if (renorm < 1.0f) { v.x *= renorm; v.y *= renorm; v.z *= renorm; v.w *= renorm; }
        }
        out[(size_t)ic * a.ncol + col] = v;
    }
}

template <int MODEL, bool HASQ0>
int dispatch(const float4* atoms, const float* cpar, const GridArgs& a, float4* out, hipStream_t st) {
    const bool dcz = (a.dC[0] == 0.f && a.dC[1] == 0.f);
    dim3 blk(256);
    if (dcz) {
        if (a.nz >= 8) {
            dim3 grd((a.ncol + 255) / 256, (a.nz + 7) / 8);
            hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 8, true>), grd, blk, 0, st, atoms, cpar, a, out);
        } else {
            dim3 grd((a.ncol + 255) / 256, (a.nz + 3) / 4);
            hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 4, true>), grd, blk, 0, st, atoms, cpar, a, out);
        }
    } else {
        dim3 grd((a.ncol + 255) / 256, (a.nz + 3) / 4);
        hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 4, false>), grd, blk, 0, st, atoms, cpar, a, out);
    }
    return check_launch("grid_project");
}

}  // namespace

int launch_grid_project(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec& g,
                        const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                        const float lvB[3], int cap, float* grid, hipStream_t stream) {
    GridArgs a{};
    a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.ncol = g.nx * g.ny;
    for (int k = 0; k < 3; ++k) {
        a.p0[k] = g.p0[k]; a.dA[k] = g.dA[k]; a.dB[k] = g.dB[k]; a.dC[k] = g.dC[k];
        a.lvA[k] = lvA ? lvA[k] : 0.f; a.lvB[k] = lvB ? lvB[k] : 0.f;
    }
    a.na = na;
    a.npbc_a = npbc_a; a.npbc_b = npbc_b;
    a.nimg_a = 2 * npbc_a + 1; a.nimg_b = 2 * npbc_b + 1;
    a.nsites = na * a.nimg_a * a.nimg_b;
    a.cap = cap;
    a.q0 = 0.f;
    a.noff = 0;
    for (int k = 0; k < 4; ++k) {
        if (Qs[k] == 0.f) continue;   // a zero-weight site adds exactly zero
        if (QZs[k] == 0.f) {
            a.q0 += Qs[k] * kCOULOMB;
        } else {
            a.qoff[a.noff] = Qs[k] * kCOULOMB;
            a.zoff[a.noff] = QZs[k];
            ++a.noff;
        }
    }
    const float4* at = reinterpret_cast<const float4*>(atoms);
    float4* o = reinterpret_cast<float4*>(grid);
    const bool hasq0 = a.q0 != 0.f;
    if (model == PPM_MODEL_LJ)
        return hasq0 ? dispatch<PPM_MODEL_LJ, true>(at, cpar, a, o, stream)
                     : dispatch<PPM_MODEL_LJ, false>(at, cpar, a, o, stream);
    return hasq0 ? dispatch<PPM_MODEL_MORSE, true>(at, cpar, a, o, stream)
                 : dispatch<PPM_MODEL_MORSE, false>(at, cpar, a, o, stream);
}

}  // namespace ppm
