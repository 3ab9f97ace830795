// ppm_f64.hip — grid projection and relaxed scan in FP64 (reference precision).
//
// Projection: the pair laws of the FP32 projection (LJ kernels/AFM.cl:738-745, Morse :747-757, point
// charges getCoulombAFM :730-736 per tip site, the |E|>100 rescale :984-985, x,y images as in
// kernels/gridFF.cl:960-996), evaluated with double-precision inputs, arithmetic and output:
// a grid whose rounding error (~1e-15 relative) is far below any finite-difference step, for
// gradient checks and reference-quality fields.  Not a hot path: one thread per voxel, the
// atom x image sites staged through LDS in tiles of 256 (64 B each), correctly rounded FP64
// division / square root / exp.
// Relaxation: relaxStrokesTilted (kernels/AFM.cl:478-562) with the PP telemetry of the direct
// path, on an FP64 grid, all arithmetic in FP64 (the FIRE and spring constants are the
// reference's float literals, widened).  One lane per scan point.  Together with the FP64
// projection it is the whole forward at reference precision, fast enough for the
// finite-difference checks of the adjoint at full image size.
#include "ppm_common.hpp"

#include <cmath>

namespace ppm {
namespace {

constexpr int kTile64 = 256;

struct GridArgs64 {
    int nx, ny, nz, ncol;
    int nimg_a, nimg_b, nsites, npbc_a, npbc_b;
    int nq;            // tip charge sites with nonzero charge
    int cap;
    double p0[3], dA[3], dB[3], dC[3], lvA[3], lvB[3];
    double qs[4];      // COULOMB * Qs of those sites
    double qz[4];      // their z offsets
};

struct Site64 {
    double x, y, z, q, c0, c1, c2, c3;
};

template <int MODEL>
__global__ __launch_bounds__(256) void k_grid_project_f64(const double* __restrict__ atoms,
                                                          const double* __restrict__ cpar, GridArgs64 a,
                                                          double* __restrict__ out) {
    __shared__ Site64 sites[kTile64];
    const int col = blockIdx.x * 256 + threadIdx.x;
    const int ic = blockIdx.y;
    const bool ok = col < a.ncol;
    const int cc = ok ? col : 0;
    const double ia = (double)(cc % a.nx), ib = (double)(cc / a.nx), fc = (double)ic;
    const double px = a.p0[0] + a.dA[0] * ia + a.dB[0] * ib + a.dC[0] * fc;
    const double py = a.p0[1] + a.dA[1] * ia + a.dB[1] * ib + a.dC[1] * fc;
    const double pz = a.p0[2] + a.dA[2] * ia + a.dB[2] * ib + a.dC[2] * fc;
    double fx = 0.0, fy = 0.0, fz = 0.0, fe = 0.0;
    const int nimg = a.nimg_a * a.nimg_b;
    for (int t0 = 0; t0 < a.nsites; t0 += kTile64) {
        __syncthreads();
        const int s = t0 + threadIdx.x;
        if (s < a.nsites) {
            const int j = s / nimg, img = s - j * nimg;
            const double ix = (double)(img % a.nimg_a - a.npbc_a), iy = (double)(img / a.nimg_a - a.npbc_b);
            Site64& S = sites[threadIdx.x];
            S.x = atoms[4 * j + 0] + ix * a.lvA[0] + iy * a.lvB[0];
            S.y = atoms[4 * j + 1] + ix * a.lvA[1] + iy * a.lvB[1];
            S.z = atoms[4 * j + 2] + ix * a.lvA[2] + iy * a.lvB[2];
            S.q = atoms[4 * j + 3];
            if (MODEL == PPM_MODEL_LJ) {
                S.c0 = cpar[2 * j]; S.c1 = cpar[2 * j + 1]; S.c2 = 0.0; S.c3 = 0.0;
            } else {
                S.c0 = cpar[4 * j]; S.c1 = cpar[4 * j + 1]; S.c2 = cpar[4 * j + 2]; S.c3 = 0.0;
            }
        }
        __syncthreads();
        const int cnt = min(kTile64, a.nsites - t0);
        for (int k = 0; k < cnt; ++k) {
            const Site64 S = sites[k];
            const double dx = px - S.x, dy = py - S.y, dz = pz - S.z;
            const double r2 = dx * dx + dy * dy + dz * dz + kR2SAFE_D;
            double fr;
            if (MODEL == PPM_MODEL_LJ) {            // c0 = c6, c1 = c12
                const double ir2 = 1.0 / r2, ir6 = ir2 * ir2 * ir2;
                fe += (S.c1 * ir6 - S.c0) * ir6;
                fr = (12.0 * S.c1 * ir6 - 6.0 * S.c0) * ir6 * ir2;
            } else {                                  // c0 = R0, c1 = E0, c2 = K
                const double r = sqrt(r2), ex = exp(S.c2 * (r - S.c0));
                fe += S.c1 * ex * (ex - 2.0);
                fr = -2.0 * S.c2 * S.c1 * ex * (ex - 1.0) / r;
            }
            fx += fr * dx; fy += fr * dy; fz += fr * dz;
            for (int t = 0; t < a.nq; ++t) {          // tip charge sites (z offset qz)
                const double dzq = dz + a.qz[t];
                const double ir2 = 1.0 / (dx * dx + dy * dy + dzq * dzq + kR2SAFE_D);
                const double e = a.qs[t] * S.q * sqrt(ir2);
                const double e2 = e * ir2;
                fx += e2 * dx; fy += e2 * dy; fz += e2 * dzq; fe += e;
            }
        }
    }
    if (!ok) return;
    if (a.cap) {
        const double renorm = 100.0 / fabs(fe);
        if (renorm < 1.0) { fx *= renorm; fy *= renorm; fz *= renorm; fe *= renorm; }
    }
    double* o = out + 4 * ((size_t)ic * a.ncol + col);
    o[0] = fx; o[1] = fy; o[2] = fz; o[3] = fe;
}

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Relax64Args {
    int nx, ny, nz_grid, n_scan, nz, max_iter, check_before, f32_anchors, damped;
    double f2conv;
    double dinv[3][4];
    double A[3], B[3], C[3], dtip, k[4], d0[4], dt0, damp0, surf;
};

// OpenCL linear-filter REPEAT addressing of one axis (OpenCL 1.2, 8.2) in FP64.
__device__ __forceinline__ void repeat64(double s, int n, int& i0, int& i1, double& a) {
    const double t = (s - floor(s)) * (double)n - 0.5;
    double fl = floor(t);
    a = t - fl;
    fl = fmin(fmax(fl, -1.0), (double)n);   // keeps non-finite coordinates in range
    int i = (int)fl, j = i + 1;
    if (i < 0) i += n;
    if (j > n - 1) j -= n;
    i0 = min(max(i, 0), n - 1);
    i1 = min(max(j, 0), n - 1);
}

// Trilinear sample in the nested order of read_imagef_trilin_ (kernels/AFM.cl:81-97).
__device__ void sample64(const double* __restrict__ g, const Relax64Args& r, D3 p, double o[4]) {
    int i[3][2];
    double w[3];
    const double pc[3] = {p.x, p.y, p.z};
    const int n[3] = {r.nx, r.ny, r.nz_grid};
    for (int ax = 0; ax < 3; ++ax) {
        const double s = pc[0] * r.dinv[ax][0] + pc[1] * r.dinv[ax][1] + pc[2] * r.dinv[ax][2] + r.dinv[ax][3];
        repeat64(s, n[ax], i[ax][0], i[ax][1], w[ax]);
    }
    const double a = w[0], b = w[1], c = w[2], ma = 1.0 - a, mb = 1.0 - b, mc = 1.0 - c;
    auto T = [&](int x, int y, int z) { return g + 4 * (((size_t)i[2][z] * r.ny + i[1][y]) * r.nx + i[0][x]); };
    for (int q = 0; q < 4; ++q) {
        const double lo = (T(0, 0, 0)[q] * ma + T(1, 0, 0)[q] * a) * mb + (T(0, 1, 0)[q] * ma + T(1, 1, 0)[q] * a) * b;
        const double hi = (T(0, 0, 1)[q] * ma + T(1, 0, 1)[q] * a) * mb + (T(0, 1, 1)[q] * ma + T(1, 1, 1)[q] * a) * b;
        o[q] = lo * mc + hi * c;
    }
}

__global__ __launch_bounds__(64) void k_relax_tilted_f64(const double* __restrict__ grid,
                                                         const double* __restrict__ pts, Relax64Args r,
                                                         double* __restrict__ FEs, double* __restrict__ PPs) {
    const int ip = blockIdx.x * 64 + threadIdx.x;
    if (ip >= r.n_scan) return;
    const D3 A = d3(r.A[0], r.A[1], r.A[2]), B = d3(r.B[0], r.B[1], r.B[2]), C = d3(r.C[0], r.C[1], r.C[2]);
    auto rot = [&](D3 v) { return d3(dot(A, v), dot(B, v), dot(C, v)); };          // world -> tip frame
    auto rotT = [&](D3 v) { return A * v.x + B * v.y + C * v.z; };               // tip frame -> world
    const D3 dTip = C * r.dtip;
    D3 tipPos = d3(pts[4 * (size_t)ip], pts[4 * (size_t)ip + 1], pts[4 * (size_t)ip + 2]);
    // the FP32 kernels' anchor: float32 position advanced by the float32 step C * dtip
    float tf[3] = {(float)tipPos.x, (float)tipPos.y, (float)tipPos.z};
    const float dtf[3] = {(float)r.C[0] * (float)r.dtip, (float)r.C[1] * (float)r.dtip, (float)r.C[2] * (float)r.dtip};
    if (r.f32_anchors) tipPos = d3(tf[0], tf[1], tf[2]);
    D3 pos = tipPos + rotT(d3(r.d0[0], r.d0[1], r.d0[2]));
    double dt = r.dt0, damp = r.damp0;
    const double dtmax = dt, dtmin = dtmax * (double)0.1f;
    for (int iz = 0; iz < r.nz; ++iz) {
        D3 v = d3(0, 0, 0);
        int it = 0;
        bool conv = false;
        double fe[4];
        for (it = 0; it < r.max_iter; ++it) {
            sample64(grid, r, pos, fe);
            // tipForce (kernels/AFM.cl:53-58) in the tip frame
            const D3 dp = rot(pos - tipPos);
            const double rr = fmax(sqrt(dot(dp, dp)), (double)1e-10f);
            const double sr = (r.k[3] * (rr - r.d0[3])) / rr;
            const D3 ft = d3((dp.x - r.d0[0]) * r.k[0], (dp.y - r.d0[1]) * r.k[1], (dp.z - r.d0[2]) * r.k[2]) + dp * sr;
            const D3 f = d3(fe[0], fe[1], fe[2]) + rotT(ft) + C * r.surf;
            const double ff = dot(f, f);
            if (r.check_before && ff < r.f2conv) { conv = true; break; }
            // update_FIRE (kernels/AFM.cl:121-157), or v *= (1 - damp) (OPT_FIRE=0, kernels/AFM.cl:536-540)
            const double vf = dot(v, f);
            if (r.damped) {
                v = v * (1.0 - damp);
            } else if (vf < 0.0) {
                v = d3(0, 0, 0);
                dt = fmax(dtmin, dt * (double)0.5f);
                damp = r.damp0;
            } else {
                const double vv = dot(v, v);
                v = v * (1.0 - damp) + f * (damp * sqrt(vv / (ff + (double)1e-8f)));
                dt = fmin(dtmax, dt * (double)1.1f);
                damp = damp * (double)0.99f;
            }
            v = v + f * dt;
            pos = pos + v * dt;
            if (!r.check_before && ff < r.f2conv) { conv = true; break; }
        }
        const int iters = conv ? it + 1 : r.max_iter;
        sample64(grid, r, pos, fe);
        const D3 fr = rot(d3(fe[0], fe[1], fe[2]));
        const size_t o = 4 * ((size_t)ip * r.nz + iz);
        FEs[o] = fr.x; FEs[o + 1] = fr.y; FEs[o + 2] = fr.z; FEs[o + 3] = fe[3];
        if (PPs) {
            const bool finite = isfinite(fe[0]) && isfinite(fe[1]) && isfinite(fe[2]) && isfinite(fe[3]) &&
                                isfinite(pos.x) && isfinite(pos.y) && isfinite(pos.z);
            PPs[o] = pos.x; PPs[o + 1] = pos.y; PPs[o + 2] = pos.z;
            PPs[o + 3] = (conv && finite) ? (double)iters : -(double)iters;
        }
        if (r.f32_anchors) {
            for (int q = 0; q < 3; ++q) tf[q] = tf[q] + dtf[q];
            tipPos = d3(tf[0], tf[1], tf[2]);
        } else {
            tipPos = tipPos + dTip;
        }
        pos = pos + dTip;
    }
}

}  // namespace

int launch_grid_project_f64(int model, const double* atoms, const double* cpar, int na, const PpmGridSpec& g,
                            const double Qs[4], const double QZs[4], int npbc_a, int npbc_b, const double lvA[3],
                            const double lvB[3], int cap, double* grid, hipStream_t stream) {
    GridArgs64 a{};
    a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.ncol = g.nx * g.ny;
    a.npbc_a = npbc_a; a.npbc_b = npbc_b;
    a.nimg_a = 2 * npbc_a + 1; a.nimg_b = 2 * npbc_b + 1;
    a.nsites = na * a.nimg_a * a.nimg_b;
    a.cap = cap;
    for (int k = 0; k < 3; ++k) {
        a.p0[k] = g.p0[k]; a.dA[k] = g.dA[k]; a.dB[k] = g.dB[k]; a.dC[k] = g.dC[k];
        a.lvA[k] = lvA ? lvA[k] : 0.0; a.lvB[k] = lvB ? lvB[k] : 0.0;
    }
    a.nq = 0;
    for (int k = 0; k < 4; ++k) {
        if (Qs[k] == 0.0) continue;   // a zero-weight site adds exactly zero
        a.qs[a.nq] = Qs[k] * (double)kCOULOMB_D;
        a.qz[a.nq] = QZs[k];
        ++a.nq;
    }
    const dim3 grd((unsigned)((a.ncol + 255) / 256), (unsigned)g.nz);
    if (model == PPM_MODEL_LJ)
        hipLaunchKernelGGL(k_grid_project_f64<PPM_MODEL_LJ>, grd, dim3(256), 0, stream, atoms, cpar, a, grid);
    else
        hipLaunchKernelGGL(k_grid_project_f64<PPM_MODEL_MORSE>, grd, dim3(256), 0, stream, atoms, cpar, a, grid);
    return check_launch("grid_project_f64");
}

int launch_relax_tilted_f64(const double* grid, const PpmGridSpec& g, const double* pts, int n_scan, int nz,
                            const PpmTip& tip, const PpmRelaxCtl& ctl, double* FEs, double* PPs,
                            hipStream_t stream) {
    Relax64Args r{};
    r.nx = g.nx; r.ny = g.ny; r.nz_grid = g.nz; r.n_scan = n_scan; r.nz = nz;
    r.max_iter = ctl.max_iter; r.check_before = ctl.check_before; r.f2conv = ctl.f2conv;
    r.f32_anchors = ctl.f32_anchors != 0;
    r.damped = ctl.damped != 0;
    for (int q = 0; q < 4; ++q) { r.dinv[0][q] = g.dinvA[q]; r.dinv[1][q] = g.dinvB[q]; r.dinv[2][q] = g.dinvC[q]; }
    for (int q = 0; q < 3; ++q) { r.A[q] = tip.tipA[q]; r.B[q] = tip.tipB[q]; r.C[q] = tip.tipC[q]; }
    r.dtip = tip.tipC[3];
    for (int q = 0; q < 4; ++q) { r.k[q] = tip.stiffness[q]; r.d0[q] = tip.dpos0[q]; }
    r.dt0 = tip.relax[0]; r.damp0 = tip.relax[1]; r.surf = tip.surfFF[0];
    hipLaunchKernelGGL(k_relax_tilted_f64, dim3((unsigned)((n_scan + 63) / 64)), dim3(64), 0, stream, grid, pts, r,
                       FEs, PPs);
    return check_launch("relax_tilted_f64");
}

}  // namespace ppm
