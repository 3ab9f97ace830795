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
//   * one workgroup = 256 columns (ia,ib) x V consecutive z voxels per thread (V = 16 or 8
//     for LJ, chosen per launch shape; 8 for Morse, 4 for a tilted dC); voxels of one thread
//     share dx, dy and dx^2+dy^2 for every site;
//   * the atom x image "sites" are expanded once per 256-site tile into LDS by the
//     workgroup (one site per thread), then read as broadcasts by every lane;
//   * the tip charge at z offset 0 shares r^2 with the pair law (one rsqrt for both);
//     charge sites with a nonzero offset are evaluated separately;
//   * stores are float4 per voxel, consecutive lanes -> consecutive x (1 KiB per wave).
#include <algorithm>

#include "ppm_grid_core.hpp"

namespace ppm {
namespace {

using namespace gc;

// Planes [klo, kend) of the grid (klo = 0, kend = nz: the whole grid), stored from plane klo on:
// a z window has the voxel values of the same planes of the whole grid, bit for bit.
template <int MODEL, bool HASQ0, int V, bool DCZ, typename ACC = float>
__global__ __launch_bounds__(256) void k_grid_project(const float4* __restrict__ atoms,
                                                      const float* __restrict__ cpar, GridArgs a,
                                                      float4* __restrict__ out, const float4* base, int klo,
                                                      int kend) {
    __shared__ Site sites[kTile + 1];
    grid_tile<MODEL, HASQ0, V, DCZ, kOffDynamic, ACC>(atoms, cpar, a, sites, blockIdx.x * 256 + threadIdx.x,
                                                      klo + blockIdx.y * V, kend, out, klo, base);
}

// ── grid -> atom adjoint ─────────────────────────────────────────────────────────────
//
// Given the cotangent g = dL/dG of every voxel (float4, same layout as the grid), accumulates
// per atom j:  dL/d(x,y,z)_j, dL/dq_j and dL/d(p0,p1)_j with (p0,p1) = (c6,c12) for LJ and
// (R0,E0) for Morse, summed over voxels and periodic images.  For one pair with
// d = v - a, s = |d|^2 + R2SAFE, F = fr(s) d, E = E(s) (dE/ds = -fr/2):
//   dL/da = -[ fr gF + (2 (gF.d) dfr/ds - gE fr) d ]
// and the parameter derivatives follow from the pair law.  Capped voxels (|E| = 100 after
// the rescale) are differentiated through the rescale: g_raw = s g - s (g.fe_raw)/E_raw e_w,
// s = 100/|E_raw|, with fe_raw recomputed for the workgroup.
// Only octets (a column x V consecutive planes) with a nonzero cotangent are evaluated: the
// probes touch a small part of the grid, and within the z blocks they reach whole columns stay
// untouched (pixel and grid spacings differ), so the octets are first listed in octet order
// (k_adj_mark, k_adj_offsets, k_adj_list) and k_grid_adjoint takes 256 listed octets and one of
// up to gridDim.z shares of the atoms per workgroup (when few octets are listed, the shares keep
// ~4096 workgroups in flight, so the last ones do not run alone on the chip; every share re-reads
// its octets' cotangent, so large lists are not split).  Each workgroup
// writes one FP64 partial sum per atom of its share (the V voxels of one site image summed in
// FP32, the rest in FP64), reduced over workgroups by k_reduce_partials in FP64.
// The list order is fixed, so the result does not depend on scheduling.

// Octet o = (z block zb, padded column col) = zb * (gridDim.x * 256) + col, the workgroup layout
// of the grid projection: flag[o] = any voxel of the octet has a nonzero cotangent, count[b] =
// flags set in workgroup b.
template <int V>
__global__ __launch_bounds__(256) void k_adj_mark(GridArgs a, const float4* __restrict__ gcot,
                                                  unsigned char* __restrict__ flag, int* __restrict__ count) {
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    const int col = blockIdx.x * 256 + threadIdx.x;
    const int ic0 = blockIdx.y * V;
    bool any = false;
    if (col < a.ncol) {
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (ic0 + k < a.nz) {
                const float4 g = gcot[(size_t)(ic0 + k) * a.ncol + col];
                any |= (g.x != 0.f) | (g.y != 0.f) | (g.z != 0.f) | (g.w != 0.f);
            }
        }
    }
    flag[(size_t)bid * 256 + threadIdx.x] = any ? 1 : 0;
    const int n = __syncthreads_count(any);
    if (threadIdx.x == 0) count[bid] = n;
}

// Exclusive offsets of the per-workgroup counts (one workgroup, each thread a contiguous run of
// counts) and their total.
__global__ __launch_bounds__(1024) void k_adj_offsets(const int* __restrict__ count, int nb, int* __restrict__ off,
                                                      int* __restrict__ total) {
    __shared__ int s[1024];
    const int per = (nb + 1023) / 1024;
    const int b0 = min(nb, (int)threadIdx.x * per), b1 = min(nb, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; ++b) sum += count[b];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = (int)threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    int run = s[threadIdx.x] - sum;
    for (int b = b0; b < b1; ++b) {
        off[b] = run;
        run += count[b];
    }
    if (threadIdx.x == 1023) *total = s[1023];
}

// list[off[b] + rank] = octet id of every flagged octet, in octet order.
__global__ __launch_bounds__(256) void k_adj_list(const unsigned char* __restrict__ flag, const int* __restrict__ off,
                                                  int* __restrict__ list) {
    __shared__ int wc[4];
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    const int o = bid * 256 + (int)threadIdx.x;
    const bool f = flag[o] != 0;
    const unsigned long long m = __ballot(f);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wc[w] = __popcll(m);
    __syncthreads();
    int base = off[bid];
    for (int i = 0; i < w; ++i) base += wc[i];
    if (f) list[base + below] = o;
}
template <int MODEL>
__device__ __forceinline__ void pair_fe_raw(float dx, float dy, float dz, float dxy2, float q, float p0, float p1,
                                            float K, const GridArgs& a, float4& fe) {
    const float r2 = fmaf(dz, dz, dxy2);
    const float ir = rsq(r2), ir2 = ir * ir;
    float fr, e;
    if (MODEL == PPM_MODEL_LJ) {
        const float ir6 = ir2 * ir2 * ir2;
        e = fmaf(p1, ir6, -p0) * ir6;
        fr = fmaf(12.f * p1, ir6, -6.f * p0) * ir6 * ir2;
    } else {
        const float r = r2 * ir;
        const float ex = __expf(K * (r - p0));
        e = p1 * ex * (ex - 2.f);
        fr = -2.f * K * p1 * ex * (ex - 1.f) * ir;
    }
    const float ec = q * a.q0 * ir;
    fr = fmaf(ec, ir2, fr);
    e += ec;
    fe.x = fmaf(fr, dx, fe.x); fe.y = fmaf(fr, dy, fe.y); fe.z = fmaf(fr, dz, fe.z); fe.w += e;
    for (int o = 0; o < a.noff; ++o) {
        const float dzo = dz + a.zoff[o];
        const float iro = rsq(fmaf(dzo, dzo, dxy2));
        const float eo = q * a.qoff[o] * iro;
        const float fro = eo * iro * iro;
        fe.x = fmaf(fro, dx, fe.x); fe.y = fmaf(fro, dy, fe.y); fe.z = fmaf(fro, dzo, fe.z); fe.w += eo;
    }
}

// Contributions of one pair to the 6 per-atom accumulators: t[0..2] += d(g.(F,E))/d(voxel - atom)
// (the caller subtracts them: the atom moves opposite to the voxel offset), t[3] += d/dq,
// t[4], t[5] += d/d(c6, c12) (LJ) or d/d(R0, E0) (Morse), only when PAR (the caller wants the
// pair-parameter gradient; without it they stay zero).  OFFS: the tip has charge sites at
// nonzero z offsets.
template <int MODEL, bool OFFS, bool PAR>
__device__ __forceinline__ void pair_adjoint(float dx, float dy, float dz, float dxy2, float q, float p0, float p1,
                                             float K, const GridArgs& a, float4 g, float t[6]) {
#pragma clang fp contract(fast)
    const float s = fmaf(dz, dz, dxy2);
    const float ir = rsq(s), ir2 = ir * ir;
    const float gd = fmaf(g.z, dz, fmaf(g.y, dy, g.x * dx));
    const float gdi = gd * ir2;
    float fr, dfr;
    if (MODEL == PPM_MODEL_LJ) {
        const float ir6 = ir2 * ir2 * ir2, ir8 = ir6 * ir2;
        const float t12 = p1 * ir6;
        fr = fmaf(12.f, t12, -6.f * p0) * ir8;
        dfr = fmaf(-84.f, t12, 24.f * p0) * (ir8 * ir2);
        if (PAR) {
            t[4] = fmaf(-ir6, fmaf(6.f, gdi, g.w), t[4]);             // d/dc6
            t[5] = fmaf(ir6 * ir6, fmaf(12.f, gdi, g.w), t[5]);       // d/dc12
        }
    } else {
        const float r = s * ir;
        const float ex = __expf(K * (r - p0));
        const float A = p1 * ex;
        fr = -2.f * K * A * (ex - 1.f) * ir;
        // d fr / d r, then / (2 r) for d/ds
        const float dfr_dr = -2.f * K * A * (K * (2.f * ex - 1.f) * ir - (ex - 1.f) * ir2);
        dfr = dfr_dr * 0.5f * ir;
        const float dE_dR0 = -2.f * K * A * (ex - 1.f);
        const float dfr_dR0 = 2.f * K * K * A * (2.f * ex - 1.f) * ir;
        if (PAR) {
            t[4] += gd * dfr_dR0 + g.w * dE_dR0;                                   // d/dR0
            t[5] += gd * (-2.f * K * ex * (ex - 1.f) * ir) + g.w * ex * (ex - 2.f); // d/dE0
        }
    }
    if (a.q0 != 0.f) {
        const float ec = q * a.q0 * ir;          // Coulomb at z offset 0 shares d with the pair law
        const float frc = ec * ir2;
        fr += frc;
        dfr = fmaf(-1.5f * frc, ir2, dfr);
        t[3] = fmaf(a.q0 * ir, gdi + g.w, t[3]);
    }
    const float coef = fmaf(2.f * gd, dfr, -g.w * fr);
    t[0] = fmaf(fr, g.x, fmaf(coef, dx, t[0]));
    t[1] = fmaf(fr, g.y, fmaf(coef, dy, t[1]));
    t[2] = fmaf(fr, g.z, fmaf(coef, dz, t[2]));
    if (OFFS) {
        for (int o = 0; o < a.noff; ++o) {
            const float dzo = dz + a.zoff[o];
            const float so = fmaf(dzo, dzo, dxy2);
            const float iro = rsq(so), iro2 = iro * iro;
            const float gdo = fmaf(g.z, dzo, fmaf(g.y, dy, g.x * dx));
            const float eo = q * a.qoff[o] * iro;
            const float fro = eo * iro2;
            const float co = 2.f * gdo * (-1.5f * fro * iro2) - g.w * fro;
            t[0] = fmaf(fro, g.x, fmaf(co, dx, t[0]));
            t[1] = fmaf(fro, g.y, fmaf(co, dy, t[1]));
            t[2] = fmaf(fro, g.z, fmaf(co, dzo, t[2]));
            t[3] = fmaf(a.qoff[o] * iro, fmaf(gdo, iro2, g.w), t[3]);
        }
    }
}

template <int MODEL, int V, bool DCZ, bool OFFS, bool PAR>
__global__ __launch_bounds__(256) void k_grid_adjoint(const float4* __restrict__ atoms, const float* __restrict__ cpar,
                                                      GridArgs a, const float4* __restrict__ grid,
                                                      const float4* __restrict__ gcot, const float4* __restrict__ base,
                                                      const int* __restrict__ list, const int* __restrict__ total,
                                                      double* __restrict__ partials, int* __restrict__ active) {
#pragma clang fp contract(fast)
    __shared__ double red[4][6];
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    const int n_list = *total;
    if (bid * 256 >= n_list) {           // the same for every thread of the workgroup
        if (threadIdx.x == 0 && blockIdx.z == 0) active[bid] = 0;
        return;
    }
    if (threadIdx.x == 0 && blockIdx.z == 0) active[bid] = 1;
    // atoms [j_begin, j_end): share blockIdx.z of n_share, n_share = the number of shares (at most
    // gridDim.z) that brings the listed workgroups to ~4096 (16 per CU)
    const int n_wg = (n_list + 255) / 256;
    const int n_share = min((int)gridDim.z, max(1, (4096 + n_wg - 1) / n_wg));
    if ((int)blockIdx.z >= n_share) return;
    const int per = (a.na + n_share - 1) / n_share;
    const int j_begin = min(a.na, (int)blockIdx.z * per), j_end = min(a.na, j_begin + per);
    const int e = bid * 256 + (int)threadIdx.x;
    const bool colok = e < n_list;
    const int oct = list[colok ? e : n_list - 1];
    const int ncolp = (int)gridDim.x * 256;
    const int zb = oct / ncolp;
    const int col = oct - zb * ncolp;     // < a.ncol: only octets holding voxels are listed
    const int ia = col % a.nx, ib = col / a.nx;
    const int ic0 = zb * V;
    float4 g[V];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        const bool ok = colok && ic0 + k < a.nz;
        g[k] = ok ? gcot[(size_t)(ic0 + k) * a.ncol + col] : make_float4(0.f, 0.f, 0.f, 0.f);
    }

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
    const int cs = MODEL == PPM_MODEL_LJ ? 2 : 4;

    // voxels clipped by the |E|>100 rescale: recompute the raw sum, chain through the rescale
    bool capped = false;
    if (a.cap) {
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const bool nzk = (g[k].x != 0.f) | (g[k].y != 0.f) | (g[k].z != 0.f) | (g[k].w != 0.f);
            if (nzk && colok && ic0 + k < a.nz) {
                if (fabsf(grid[(size_t)(ic0 + k) * a.ncol + col].w) >= 99.999f) capped = true;
            }
        }
    }
    if (__syncthreads_or(capped)) {
        float4 fe[V];
#pragma unroll
        for (int k = 0; k < V; ++k) fe[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int j = 0; j < a.na; ++j) {
            const float4 at = atoms[j];
            const float p0 = cpar[j * cs], p1 = cpar[j * cs + 1], K = MODEL == PPM_MODEL_LJ ? 0.f : cpar[j * cs + 2];
            for (int iy = -a.npbc_b; iy <= a.npbc_b; ++iy)
                for (int ix = -a.npbc_a; ix <= a.npbc_a; ++ix) {
                    const float sx = at.x + ix * a.lvA[0] + iy * a.lvB[0];
                    const float sy = at.y + ix * a.lvA[1] + iy * a.lvB[1];
                    const float sz = at.z + ix * a.lvA[2] + iy * a.lvB[2];
#pragma unroll
                    for (int k = 0; k < V; ++k) {
                        const int kk = DCZ ? 0 : k;
                        const float dx = px[kk] - sx, dy = py[kk] - sy;
                        pair_fe_raw<MODEL>(dx, dy, pz[k] - sz, fmaf(dx, dx, fmaf(dy, dy, kR2SAFE)), at.w, p0, p1, K,
                                           a, fe[k]);
                    }
                }
        }
        if (base) {            // the raw sum includes the base grid's atoms
#pragma unroll
            for (int k = 0; k < V; ++k) {
                if (colok && ic0 + k < a.nz) {
                    const float4 b = base[(size_t)(ic0 + k) * a.ncol + col];
                    fe[k].x += b.x; fe[k].y += b.y; fe[k].z += b.z; fe[k].w += b.w;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if (fabsf(fe[k].w) > 100.f) {
                const float sc = 100.f / fabsf(fe[k].w);
                const float gdf = g[k].x * fe[k].x + g[k].y * fe[k].y + g[k].z * fe[k].z + g[k].w * fe[k].w;
                g[k].x *= sc; g[k].y *= sc; g[k].z *= sc;
                g[k].w = sc * g[k].w - sc * gdf / fe[k].w;
            }
        }
    }

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NC = PAR ? 6 : 4;      // accumulated columns: xyz, q (+ the pair parameters)
    for (int j = j_begin; j < j_end; ++j) {
        const float4 at = atoms[j];
        const float p0 = cpar[j * cs], p1 = cpar[j * cs + 1], K = MODEL == PPM_MODEL_LJ ? 0.f : cpar[j * cs + 2];
        // The V voxels of one image are summed in FP32, everything above in FP64: the grid
        // cotangent alternates in sign along z (it is a Δf adjoint), so contractions with the
        // long-range charge field cancel by ~1e3 and an FP32 sum over the workgroup would lose
        // ~1e-3 of the charge gradient.
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int iy = -a.npbc_b; iy <= a.npbc_b; ++iy)
            for (int ix = -a.npbc_a; ix <= a.npbc_a; ++ix) {
                const float sx = at.x + ix * a.lvA[0] + iy * a.lvB[0];
                const float sy = at.y + ix * a.lvA[1] + iy * a.lvB[1];
                const float sz = at.z + ix * a.lvA[2] + iy * a.lvB[2];
                float t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < V; ++k) {
                    const int kk = DCZ ? 0 : k;
                    const float dx = px[kk] - sx, dy = py[kk] - sy;
                    pair_adjoint<MODEL, OFFS, PAR>(dx, dy, pz[k] - sz, fmaf(dx, dx, fmaf(dy, dy, kR2SAFE)), at.w,
                                                   p0, p1, K, a, g[k], t);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[c] -= (double)t[c];
#pragma unroll
                for (int c = 3; c < NC; ++c) acc[c] += (double)t[c];
            }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double v = acc[c];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[wave][c] = v;
        }
        __syncthreads();
        if (threadIdx.x < 6)
            partials[((size_t)bid * a.na + j) * 6 + threadIdx.x] =
                (int)threadIdx.x < NC
                    ? red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x]
                    : 0.0;
        __syncthreads();
    }
}

// grad[j*6 + c] = sum over active workgroups b of partials[(b*na + j)*6 + c], in FP64.
__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ partials,
                                                         const int* __restrict__ active, int nblocks, int na,
                                                         double* __restrict__ grad) {
    __shared__ double red[256];
    const int j = blockIdx.x / 6, c = blockIdx.x - (blockIdx.x / 6) * 6;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256)
        if (active[b]) acc += partials[((size_t)b * na + j) * 6 + c];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[j * 6 + c] = red[0];
}

// 16 or 8 voxels per thread for an LJ launch of ncol columns x nzw planes.  Per voxel, 16 is ~8 %
// cheaper once every compute unit holds several workgroups, but its workgroups take twice as long,
// pad z to a multiple of 16, and, one per compute unit (one wave per SIMD), run latency-bound.
// With k = ceil(workgroups / compute units), the busiest unit's share, the launch time of either
// form, measured on MI355X over 128^2..480^2 columns x 16..256 planes, fits (least squares)
//   16 voxels: 61.0 k + 29 (+ 12 when k = 1),   8 voxels: 33.4 k + 18
// in units of 10 us for 1800 sites (200 atoms x 9 images); the ratio of the two forms does not
// depend on the site count (20..600 atoms measured), so the smaller prediction picks the form
// (within 1.4 % of the faster one on every measured shape).  The two forms sum each voxel's pair
// terms in the same order: their grids are bit-identical.
bool lj_16_voxels(int ncol, int nzw) {
    const long long cus = device_cus();
    const long long cb = (ncol + 255) / 256;
    const long long k16 = (cb * ((nzw + 15) / 16) + cus - 1) / cus;
    const long long k8 = (cb * ((nzw + 7) / 8) + cus - 1) / cus;
    return 610 * k16 + 290 + (k16 == 1 ? 120 : 0) <= 334 * k8 + 180;
}

template <int MODEL, bool HASQ0>
int dispatch(const float4* atoms, const float* cpar, const GridArgs& a, float4* out, const float4* base,
             bool acc64, hipStream_t st, int klo, int nzw) {
    const bool dcz = (a.dC[0] == 0.f && a.dC[1] == 0.f);
    const int kend = klo + nzw;
    dim3 blk(256);
    if (acc64) {          // FP64 sums: 8 voxels per thread (16 would need 128 VGPRs of sums alone)
        const int V = dcz ? 8 : 4;
        dim3 grd((a.ncol + 255) / 256, (nzw + V - 1) / V);
        if (dcz)
            hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 8, true, double>), grd, blk, 0, st, atoms, cpar, a, out,
                               base, klo, kend);
        else
            hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 4, false, double>), grd, blk, 0, st, atoms, cpar, a,
                               out, base, klo, kend);
        return check_launch("grid_project_acc64");
    }
    if (dcz) {
        if (MODEL == PPM_MODEL_LJ && nzw >= 16 && lj_16_voxels(a.ncol, nzw)) {   // 16 on Morse: +2 %
            dim3 grd((a.ncol + 255) / 256, (nzw + 15) / 16);
            hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 16, true>), grd, blk, 0, st, atoms, cpar, a, out, base,
                               klo, kend);
        } else if (nzw >= 8) {
            dim3 grd((a.ncol + 255) / 256, (nzw + 7) / 8);
            hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 8, true>), grd, blk, 0, st, atoms, cpar, a, out, base,
                               klo, kend);
        } else {
            dim3 grd((a.ncol + 255) / 256, (nzw + 3) / 4);
            hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 4, true>), grd, blk, 0, st, atoms, cpar, a, out, base,
                               klo, kend);
        }
    } else {
        dim3 grd((a.ncol + 255) / 256, (nzw + 3) / 4);
        hipLaunchKernelGGL((k_grid_project<MODEL, HASQ0, 4, false>), grd, blk, 0, st, atoms, cpar, a, out, base, klo,
                           kend);
    }
    return check_launch("grid_project");
}

}  // namespace

static GridArgs make_args(int na, const PpmGridSpec& g, const float Qs[4], const float QZs[4], int npbc_a,
                          int npbc_b, const float lvA[3], const float lvB[3], int cap) {
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
    return a;
}

// workgroup grid of both the projection and its adjoint
static dim3 grid_blocks(const PpmGridSpec& g, int V) {
    return dim3((unsigned)((g.nx * g.ny + 255) / 256), (unsigned)((g.nz + V - 1) / V));
}

static int adjoint_V(const PpmGridSpec& g) { return (g.dC[0] == 0.f && g.dC[1] == 0.f) ? 8 : 4; }

int launch_grid_project(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec& g,
                        const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                        const float lvB[3], int cap, const float* base, float* grid, bool acc64,
                        hipStream_t stream, int klo, int nzw) {
    const GridArgs a = make_args(na, g, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap);
    const float4* at = reinterpret_cast<const float4*>(atoms);
    float4* o = reinterpret_cast<float4*>(grid);
    const float4* b = reinterpret_cast<const float4*>(base);
    const bool hasq0 = a.q0 != 0.f;
    if (nzw <= 0) { klo = 0; nzw = g.nz; }
    if (model == PPM_MODEL_LJ)
        return hasq0 ? dispatch<PPM_MODEL_LJ, true>(at, cpar, a, o, b, acc64, stream, klo, nzw)
                     : dispatch<PPM_MODEL_LJ, false>(at, cpar, a, o, b, acc64, stream, klo, nzw);
    return hasq0 ? dispatch<PPM_MODEL_MORSE, true>(at, cpar, a, o, b, acc64, stream, klo, nzw)
                 : dispatch<PPM_MODEL_MORSE, false>(at, cpar, a, o, b, acc64, stream, klo, nzw);
}

// Workspace of the grid adjoint: per-workgroup FP64 partial rows, active flags, octet flags,
// per-workgroup octet counts and offsets, the octet count and the octet list, each 256-B aligned.
struct AdjWorkspace {
    double* partials;
    int* active;
    unsigned char* flag;
    int *count, *off, *total, *list;
};

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

static size_t adjoint_layout(int na, const PpmGridSpec& g, void* base, AdjWorkspace* w) {
    const dim3 b = grid_blocks(g, adjoint_V(g));
    const size_t nb = (size_t)b.x * b.y, noct = nb * 256;
    const size_t sizes[7] = {nb * (size_t)na * 6 * sizeof(double), nb * sizeof(int), noct, nb * sizeof(int),
                             nb * sizeof(int), sizeof(int), noct * sizeof(int)};
    size_t off[7], at = 0;
    for (int i = 0; i < 7; ++i) {
        off[i] = at;
        at += align256(sizes[i]);
    }
    if (w) {
        char* p = reinterpret_cast<char*>(base);
        w->partials = reinterpret_cast<double*>(p + off[0]);
        w->active = reinterpret_cast<int*>(p + off[1]);
        w->flag = reinterpret_cast<unsigned char*>(p + off[2]);
        w->count = reinterpret_cast<int*>(p + off[3]);
        w->off = reinterpret_cast<int*>(p + off[4]);
        w->total = reinterpret_cast<int*>(p + off[5]);
        w->list = reinterpret_cast<int*>(p + off[6]);
    }
    return at + 256;
}

size_t grid_adjoint_workspace(int na, const PpmGridSpec& g) { return adjoint_layout(na, g, nullptr, nullptr); }

int launch_grid_adjoint(int model, const float* atoms, const float* cpar, int na, const PpmGridSpec& g,
                        const float Qs[4], const float QZs[4], int npbc_a, int npbc_b, const float lvA[3],
                        const float lvB[3], int cap, const float* grid, const float* grid_cot, const float* base,
                        void* workspace, int params, double* grad, hipStream_t stream) {
    const GridArgs a = make_args(na, g, Qs, QZs, npbc_a, npbc_b, lvA, lvB, cap);
    const int V = adjoint_V(g);
    const dim3 blocks = grid_blocks(g, V);
    const size_t nb = (size_t)blocks.x * blocks.y;
    // atom shares per listed workgroup: 4 while each share keeps >= 8 atoms
    const dim3 adj_blocks(blocks.x, blocks.y, (unsigned)std::max(1, std::min(4, na / 8)));
    if (nb * 256 > (size_t)INT32_MAX) {
        set_error("grid_adjoint: %zu voxel octets exceed the int range", nb * 256);
        return PPM_ERR_INVALID_ARG;
    }
    AdjWorkspace w;
    adjoint_layout(na, g, workspace, &w);
    double* partials = w.partials;
    int* active = w.active;
    const float4* at = reinterpret_cast<const float4*>(atoms);
    const float4* G = reinterpret_cast<const float4*>(grid);
    const float4* gc = reinterpret_cast<const float4*>(grid_cot);
    const float4* B = reinterpret_cast<const float4*>(base);
    const bool offs = a.noff > 0;
    if (V == 8) hipLaunchKernelGGL(k_adj_mark<8>, blocks, dim3(256), 0, stream, a, gc, w.flag, w.count);
    else hipLaunchKernelGGL(k_adj_mark<4>, blocks, dim3(256), 0, stream, a, gc, w.flag, w.count);
    hipLaunchKernelGGL(k_adj_offsets, dim3(1), dim3(1024), 0, stream, w.count, (int)nb, w.off, w.total);
    hipLaunchKernelGGL(k_adj_list, blocks, dim3(256), 0, stream, w.flag, w.off, w.list);
    if (int r = check_launch("grid_adjoint_list")) return r;
#define PPM_ADJ1(M, VV, DCZ, OF, PA)                                                                         \
    hipLaunchKernelGGL((k_grid_adjoint<M, VV, DCZ, OF, PA>), adj_blocks, dim3(256), 0, stream, at, cpar, a, G, gc, \
                       B, w.list, w.total, partials, active)
#define PPM_ADJ(M, VV, DCZ, OF)                                                                              \
    do {                                                                                                     \
        if (params) PPM_ADJ1(M, VV, DCZ, OF, true);                                                         \
        else PPM_ADJ1(M, VV, DCZ, OF, false);                                                               \
    } while (0)
    if (model == PPM_MODEL_LJ) {
        if (V == 8) { if (offs) PPM_ADJ(PPM_MODEL_LJ, 8, true, true); else PPM_ADJ(PPM_MODEL_LJ, 8, true, false); }
        else { if (offs) PPM_ADJ(PPM_MODEL_LJ, 4, false, true); else PPM_ADJ(PPM_MODEL_LJ, 4, false, false); }
    } else {
        if (V == 8) { if (offs) PPM_ADJ(PPM_MODEL_MORSE, 8, true, true); else PPM_ADJ(PPM_MODEL_MORSE, 8, true, false); }
        else { if (offs) PPM_ADJ(PPM_MODEL_MORSE, 4, false, true); else PPM_ADJ(PPM_MODEL_MORSE, 4, false, false); }
    }
#undef PPM_ADJ
#undef PPM_ADJ1
    if (int r = check_launch("grid_adjoint")) return r;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)(na * 6)), dim3(256), 0, stream, partials, active, (int)nb,
                       na, grad);
    return check_launch("grid_adjoint_reduce");
}

}  // namespace ppm
