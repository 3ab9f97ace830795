// ppm_grid_core.hpp — device code of the force-field grid projection shared by the plain
// and the batched kernels (semantics and reference citations in ppm_grid.hip).  Built with
// -ffp-contract=off; every fused multiply-add is written explicitly.
#pragma once
#pragma clang fp contract(off)
#include "ppm_common.hpp"

namespace ppm {
namespace gc {

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

constexpr int kOffNone = 0, kOffDynamic = 1, kOffStatic = 2;

// Coulomb term of one tip charge site at z offset zo (charge fraction qf) on the V voxels.
// Accumulation of one pair term into a voxel sum: FP32 (fused multiply-add) or, for ACC =
// double, the FP32 term added in FP64 (the product of two floats is exact in double).
__device__ __forceinline__ void acc_fma(float& s, float a, float b) { s = fmaf(a, b, s); }
__device__ __forceinline__ void acc_fma(double& s, float a, float b) { s = fma((double)a, (double)b, s); }
__device__ __forceinline__ void acc_add(float& s, float t) { s += t; }
__device__ __forceinline__ void acc_add(double& s, float t) { s += (double)t; }

template <int V, bool DCZ, typename ACC>
__device__ __forceinline__ void offset_site(const Site& S, float qf, float zo, const float (&pz)[V],
                                            const float (&dxs)[DCZ ? 1 : V], const float (&dys)[DCZ ? 1 : V],
                                            const float (&dxy2)[DCZ ? 1 : V], ACC (&fx)[V], ACC (&fy)[V],
                                            ACC (&fz)[V], ACC (&fe)[V]) {
    const float qo = S.q * qf;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int kk = DCZ ? 0 : k;
        const float dzo = (pz[k] - S.z) + zo;
        const float iro = rsq(fmaf(dzo, dzo, dxy2[kk]));
        const float eo = qo * iro;
        const float fro = eo * iro * iro;
        acc_fma(fx[k], fro, dxs[kk]);
        acc_fma(fy[k], fro, dys[kk]);
        acc_fma(fz[k], fro, dzo);
        acc_add(fe[k], eo);
    }
}

// One workgroup tile: 256 columns starting at column `col - threadIdx.x`, V planes from ic0.
// Voxels with ic < ic_end are stored at out[(ic - out_klo)*ncol + col].  Every thread of the
// workgroup must call it (it synchronizes on the shared site tile); `sites` holds kTile + 1
// entries.
// `base` (optional, never the same buffer as `out`): a raw (uncapped) grid of other atoms added
// to every voxel before the |E|>100 rescale.
// V voxels per thread along c (z); DCZ: dC has no x,y component (every grid set-up of
// spammm/SPM/AFM.py:772-893 has dC = (0,0,L/nz)), so all V voxels share x,y.
// OFF: how the tip's offset charge sites (a.noff of them) are looped over.  kOffDynamic: run-time
// bound (GridArgs in kernel-argument memory, as in the plain launch); kOffStatic: compile-time
// bound with an early exit (GridArgs in registers, as in the batched kernel, where a dynamic
// index would put qoff/zoff in scratch); kOffNone: no offset sites (a.noff == 0).  The static
// form schedules the pair loop ~20 % slower than the other two on gfx950.
// ACC: type of the per-voxel sums.  float: the FP32 sum in site order.  double: every pair term
// is evaluated in FP32 exactly as for float and added in FP64, so the stored voxel carries only
// the terms' own rounding, not the growth of an FP32 running sum over thousands of terms that
// cancel (a field at a few A above a neutral molecule is ~1e-2 of its largest terms).
template <int MODEL, bool HASQ0, int V, bool DCZ, int OFF = kOffDynamic, typename ACC = float>
__device__ __forceinline__ void grid_tile(const float4* __restrict__ atoms, const float* __restrict__ cpar,
                                          const GridArgs& a, Site* sites, int col, int ic0, int ic_end,
                                          float4* __restrict__ out, int out_klo, const float4* base = nullptr) {
    const bool colok = col < a.ncol;
    const int cc = colok ? col : a.ncol - 1;
    const int ia = cc % a.nx, ib = cc / a.nx;

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
    ACC fx[V], fy[V], fz[V], fe[V];
#pragma unroll
    for (int k = 0; k < V; ++k) { fx[k] = 0; fy[k] = 0; fz[k] = 0; fe[k] = 0; }

    for (int t0 = 0; t0 < a.nsites; t0 += kTile) {
        __syncthreads();
        const int s = t0 + threadIdx.x;
        if (s < a.nsites) load_site<MODEL>(atoms, cpar, a, s, sites[threadIdx.x]);
        __syncthreads();
        const int cnt = min(kTile, a.nsites - t0);
        // the next site is read one iteration ahead (sites[] has kTile + 1 entries; the read
        // of entry cnt is never used), so the LDS latency overlaps the current site's math
        Site Sn = sites[0];
        for (int j = 0; j < cnt; ++j) {
            const Site S = Sn;
            Sn = sites[j + 1];
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
                float fr;
                if constexpr (sizeof(ACC) == sizeof(double)) {
                    float e;                                  // the pair's energy term
                    if (MODEL == PPM_MODEL_LJ) {
                        const float ir6 = ir2 * ir2 * ir2;
                        if (HASQ0) {
                            const float ec = qq0 * ir;
                            e = fmaf(fmaf(S.c1, ir6, -S.c0), ir6, ec);
                            fr = fmaf(fmaf(S.c3, ir6, -S.c2), ir6, ec) * ir2;
                        } else {
                            e = fmaf(S.c1, ir6, -S.c0) * ir6;
                            fr = fmaf(S.c3, ir6, -S.c2) * ir6 * ir2;
                        }
                    } else {
                        const float r = r2 * ir;
                        const float ex = ex2(fmaf(S.c2, r, S.c0));
                        const float Ae = S.c1 * ex;
                        e = (ex - 2.0f) * Ae;
                        fr = S.c3 * (Ae * (ex - 1.0f)) * ir;
                        if (HASQ0) {
                            const float ec = qq0 * ir;
                            fr = fmaf(ec, ir2, fr);
                            e += ec;
                        }
                    }
                    acc_add(fe[k], e);
                } else if (MODEL == PPM_MODEL_LJ && HASQ0) {
                    // LJ and the charge term share the final 1/r^2: 9 VALU instead of 10
                    const float ir6 = ir2 * ir2 * ir2;
                    const float ec = qq0 * ir;
                    fe[k] = fmaf(fmaf(S.c1, ir6, -S.c0), ir6, fe[k] + ec);  // E += ec + (c12 ir6 - c6) ir6
                    fr = fmaf(fmaf(S.c3, ir6, -S.c2), ir6, ec) * ir2;       // fr = ((c3 ir6 - c2) ir6 + ec) / r^2
                } else if (MODEL == PPM_MODEL_LJ) {
                    const float ir6 = ir2 * ir2 * ir2;
                    fe[k] = fmaf(fmaf(S.c1, ir6, -S.c0), ir6, fe[k]);      // E = (c12 ir6 - c6) ir6
                    fr = fmaf(S.c3, ir6, -S.c2) * ir6 * ir2;
                } else {
                    const float r = r2 * ir;
                    const float ex = ex2(fmaf(S.c2, r, S.c0));
                    const float Ae = S.c1 * ex;
                    fe[k] = fmaf(ex - 2.0f, Ae, fe[k]);                      // E = E0 ex (ex - 2)
                    fr = S.c3 * (Ae * (ex - 1.0f)) * ir;
                }
                if (sizeof(ACC) == sizeof(float) && HASQ0 && MODEL != PPM_MODEL_LJ) {
                    const float ec = qq0 * ir;
                    fr = fmaf(ec, ir2, fr);
                    fe[k] += (float)ec;
                }
                acc_fma(fx[k], fr, dx);
                acc_fma(fy[k], fr, dy);
                acc_fma(fz[k], fr, dz);
            }
            // tip charge sites with a nonzero z offset (none for a single-charge probe)
            if constexpr (OFF == kOffStatic) {
#pragma unroll
                for (int o = 0; o < 4; ++o) {
                    if (o >= a.noff) break;
                    offset_site<V, DCZ, ACC>(S, a.qoff[o], a.zoff[o], pz, dxs, dys, dxy2, fx, fy, fz, fe);
                }
            } else if constexpr (OFF == kOffDynamic) {
                for (int o = 0; o < a.noff; ++o)
                    offset_site<V, DCZ, ACC>(S, a.qoff[o], a.zoff[o], pz, dxs, dys, dxy2, fx, fy, fz, fe);
            }
        }
    }
    if (!colok) return;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const int ic = ic0 + k;
        if (ic >= ic_end) break;
        float4 v = make_float4((float)fx[k], (float)fy[k], (float)fz[k], (float)fe[k]);
        if (base) {            // raw sum of other atoms (e.g. a substrate), same layout as out
            const float4 b = base[(size_t)(ic - out_klo) * a.ncol + col];
            v.x += b.x; v.y += b.y; v.z += b.z; v.w += b.w;
        }
        if (a.cap) {
            const float renorm = 100.0f / fabsf(v.w);
            if (renorm < 1.0f) { v.x *= renorm; v.y *= renorm; v.z *= renorm; v.w *= renorm; }
        }
        out[(size_t)(ic - out_klo) * a.ncol + col] = v;
    }
}

}  // namespace gc
}  // namespace ppm
