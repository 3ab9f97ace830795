// ppm_direct.hip — gridless direct Morse + point-charge probe relaxation and its
// implicit-equilibrium adjoint (gfx950).
//
// Forward (relaxStrokesTiltedMorseDirect, kernels/AFM.cl:1419-1511): the relaxation of
// relaxStrokesTilted with the trilinear grid sample replaced by the direct, uncapped pair sum
// over all atoms (getMorse + getCoulombAFM per tip charge site, kernels/AFM.cl:1391-1407),
// convergence tested before the update.  A scan point is relaxed by a group of G lanes
// (G = 1..64, chosen so that a launch holds about 4 waves per SIMD: a 64x64 scan is only 64
// waves at one lane per point): each lane sums a strided share of the atoms (held in LDS,
// <= 128), a butterfly gives every lane of the group the total, and the group runs the FIRE
// logic in lockstep; groups advance through their z slices independently.  Built with
// -ffp-contract=off: the FP32 operations follow the reference formulas as written (with G = 1
// the atoms are also summed in the reference's order).
//
// Adjoint (morseDirectStateAdjoint + morseDirectParamPartials + reduceMorseDirectParamPartials,
// kernels/AFM.cl:1711-2029) in two launches:
//   k_direct_adjoint  one workgroup = 256 consecutive (scan, z) states.  Each thread evaluates
//                     its state in FP64 (sample force, energy Hessian, tip spring Jacobian),
//                     applies the reference's gates, solves J lambda = b and keeps
//                     (q*, w = R^T u - lambda, u.w) in LDS; then the threads turn to atoms
//                     and sum the per-atom partials over the block's states from LDS, in
//                     FP64, and write one partial row per (block, atom).
//   k_direct_reduce   one workgroup per atom sums the block partials (FP64, fixed order).
// No atomics: the gradient is deterministic.
#include "ppm_common.hpp"
#include "ppm_relax_core.hpp"

namespace ppm {
namespace {

constexpr int kMaxAtoms = PPM_DIRECT_MAX_ATOMS;
constexpr int kAdjBlock = 256;

struct DirectArgs {
    int na, nsite;
    float Qs[4], QZs[4];     // non-zero tip charge sites only, Qs scaled by COULOMB_CONST
};

__device__ __forceinline__ void load_atoms(const float4* __restrict__ atoms, const float4* __restrict__ cms, int na,
                                           float4* A, float4* M) {
    for (int i = threadIdx.x; i < na; i += blockDim.x) {
        A[i] = atoms[i];
        M[i] = cms[i];
    }
    __syncthreads();
}

// Pair sum of one lane of a G-lane group: atoms sub, sub+G, sub+2G, ... then a butterfly over
// the group, after which every lane of the group holds the same total (FP32 addition is
// commutative, so both partners of each exchange compute the identical sum).
template <int G>
__device__ __forceinline__ float4 direct_fe_group(const float4* __restrict__ A, const float4* __restrict__ M,
                                                  const DirectArgs& d, int sub, float x, float y, float z) {
    float4 fe = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int j = sub; j < d.na; j += G) {
        const float4 a = A[j];
        const float4 m = M[j];
        const float dx = x - a.x, dy = y - a.y, dz = z - a.z;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz + kR2SAFE);
        const float ex = expf(m.z * (r - m.x));
        const float E = m.y * ex * (ex - 2.0f);
        const float fr = -m.y * ex * (ex - 1.0f) * 2.0f * m.z;
        const float fq = fr / r;
        fe.x += dx * fq; fe.y += dy * fq; fe.z += dz * fq; fe.w += E;
        for (int k = 0; k < d.nsite; ++k) {
            const float kx = x - a.x, ky = y - a.y, kz = (z + d.QZs[k]) - a.z;
            const float ir2 = 1.0f / (kx * kx + ky * ky + kz * kz + kR2SAFE);
            const float Ec = a.w * sqrtf(ir2);
            const float e2 = Ec * ir2;
            fe.x += (kx * e2) * d.Qs[k]; fe.y += (ky * e2) * d.Qs[k]; fe.z += (kz * e2) * d.Qs[k];
            fe.w += Ec * d.Qs[k];
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        fe.x += __shfl_xor(fe.x, off, G);
        fe.y += __shfl_xor(fe.y, off, G);
        fe.z += __shfl_xor(fe.z, off, G);
        fe.w += __shfl_xor(fe.w, off, G);
    }
    return fe;
}

// G lanes per scan point (G = 1: one lane per point, the atoms summed in order).  The group's
// lanes share every branch (same data), so the shuffles only read active lanes.
template <int G>
__global__ __launch_bounds__(256) void k_direct_relax(const float4* __restrict__ atoms, const float4* __restrict__ cms,
                                                      DirectArgs d, const float4* __restrict__ pts, rc::TiltArgs p,
                                                      float4* __restrict__ FEs, float4* __restrict__ PPs) {
    __shared__ float4 A[kMaxAtoms], M[kMaxAtoms];
    load_atoms(atoms, cms, d.na, A, M);        // every lane takes part before any leaves
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ip = t / G, sub = t - ip * G;
    if (ip >= p.n_scan) return;                // whole groups leave together
    const float4 P = pts[ip];
    auto smp = [&](float x, float y, float z) { return direct_fe_group<G>(A, M, d, sub, x, y, z); };
    const size_t o = (size_t)ip * p.nz;
    rc::relax_tilted_lane_src<G>(smp, mk3(P.x, P.y, P.z), p, FEs + o, PPs ? PPs + o : nullptr, sub);
}

// Lanes per scan point: enough groups for about 4 waves per SIMD (1024 SIMDs), at most one
// lane per atom, a power of two in [1, 64].
int direct_group(int n_scan, int na) {
    const long long want = 4LL * 1024 * 64;
    int g = 1;
    while (g < 64 && (long long)n_scan * g * 2 <= want && g < na) g *= 2;
    return g;
}

// ── adjoint ──────────────────────────────────────────────────────────────────────────

struct AdjArgs {
    int na, nsite, n_scan, nz;
    double Qs[4], QZs[4];
    double R[3][3];          // rows tipA, tipB, tipC
    double dtip;
    double k[4], d0[4];
    double surf;
};

// Per-atom FP64 pair quantities at probe position q (kernels/AFM.cl:1546-1614, :1873-1940).
struct PairQ {
    double F[3], H[6];       // force on the probe; Hessian (xx, xy, xz, yy, yz, zz)
    double dFR0, dFE0;       // dF/dR0 = d * dFR0, dF/dE0 = d * dFE0 (d = q - a)
    double dER0, dEE0;
    double dFQ[3], dEQ;
    double d[3];
};

__device__ __forceinline__ void pair_q(const double q[3], float4 a, float4 m, const AdjArgs& g, PairQ& o) {
    o.d[0] = q[0] - (double)a.x; o.d[1] = q[1] - (double)a.y; o.d[2] = q[2] - (double)a.z;
    const double* d = o.d;
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + (double)kR2SAFE;
    const double r = sqrt(r2), ir = 1.0 / r;
    const double R0 = m.x, E0 = m.y, K = m.z;
    const double s = exp(K * (r - R0));
    const double Ep = 2.0 * K * E0 * s * (s - 1.0);
    const double Epp = 2.0 * K * K * E0 * s * (2.0 * s - 1.0);
    const double co = (Epp - Ep * ir) / r2, cd = Ep * ir;
    for (int i = 0; i < 3; ++i) o.F[i] = -d[i] * Ep * ir;
    o.H[0] = co * d[0] * d[0] + cd; o.H[1] = co * d[0] * d[1]; o.H[2] = co * d[0] * d[2];
    o.H[3] = co * d[1] * d[1] + cd; o.H[4] = co * d[1] * d[2]; o.H[5] = co * d[2] * d[2] + cd;
    o.dFR0 = 2.0 * K * K * E0 * s * (2.0 * s - 1.0) * ir;
    o.dFE0 = -2.0 * K * s * (s - 1.0) * ir;
    o.dER0 = -2.0 * K * E0 * s * (s - 1.0);
    o.dEE0 = s * (s - 2.0);
    o.dFQ[0] = o.dFQ[1] = o.dFQ[2] = 0.0;
    o.dEQ = 0.0;
    for (int k = 0; k < g.nsite; ++k) {
        const double kz = d[2] + g.QZs[k];
        const double rk2 = d[0] * d[0] + d[1] * d[1] + kz * kz + (double)kR2SAFE;
        const double rk = sqrt(rk2), irk = 1.0 / rk, irk3 = irk / rk2, irk5 = irk3 / rk2;
        const double C = (double)a.w * g.Qs[k];
        const double dk[3] = {d[0], d[1], kz};
        for (int i = 0; i < 3; ++i) {
            o.F[i] += dk[i] * C * irk3;
            o.dFQ[i] += dk[i] * g.Qs[k] * irk3;
        }
        o.dEQ += g.Qs[k] * irk;
        const double c3 = 3.0 * C * irk5, c1 = C * irk3;
        o.H[0] += c3 * dk[0] * dk[0] - c1; o.H[1] += c3 * dk[0] * dk[1]; o.H[2] += c3 * dk[0] * dk[2];
        o.H[3] += c3 * dk[1] * dk[1] - c1; o.H[4] += c3 * dk[1] * dk[2]; o.H[5] += c3 * dk[2] * dk[2] - c1;
    }
}

__device__ __forceinline__ double eigmin_sym3(const double S[3][3]) {
    // O. K. Smith, Commun. ACM 4(4):168 (1961)
    const double p1 = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
    if (p1 == 0.0) return fmin(S[0][0], fmin(S[1][1], S[2][2]));
    const double q = (S[0][0] + S[1][1] + S[2][2]) / 3.0;
    const double a = S[0][0] - q, b = S[1][1] - q, c = S[2][2] - q;
    const double p = sqrt((a * a + b * b + c * c + 2.0 * p1) / 6.0);
    const double ip = 1.0 / p;
    const double B00 = a * ip, B11 = b * ip, B22 = c * ip, B01 = S[0][1] * ip, B02 = S[0][2] * ip, B12 = S[1][2] * ip;
    const double detB = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
    const double rr = fmin(fmax(0.5 * detB, -1.0), 1.0);
    return q + 2.0 * p * cos(acos(rr) / 3.0 + 2.0943951023931957);
}

struct StateL {              // what phase 2 needs of a state, in LDS
    double q[3], w[3], uw;
    int ok;
};

__global__ __launch_bounds__(kAdjBlock) void k_direct_adjoint(const float4* __restrict__ atoms,
                                                              const float4* __restrict__ cms,
                                                              const float4* __restrict__ pts, AdjArgs g,
                                                              const float4* __restrict__ PPs,
                                                              const float4* __restrict__ U,
                                                              float4* __restrict__ lambdas, float4* __restrict__ diag,
                                                              double* __restrict__ part) {
    __shared__ float4 A[kMaxAtoms], M[kMaxAtoms];
    __shared__ StateL SL[kAdjBlock];
    __shared__ double red[kAdjBlock][6];
    load_atoms(atoms, cms, g.na, A, M);
    const size_t n_state = (size_t)g.n_scan * g.nz;
    const size_t s = (size_t)blockIdx.x * kAdjBlock + threadIdx.x;
    StateL me;
    me.ok = 0;
    if (s < n_state) {
        const int ip = (int)(s / g.nz), iz = (int)(s - (size_t)ip * g.nz);
        const float4 pp = PPs[s], u = U[s];
        float4 dg = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 lam = make_float4(0.f, 0.f, 0.f, 0.f);
        int status = 0;
        if (!(isfinite(pp.x) && isfinite(pp.y) && isfinite(pp.z) && isfinite(pp.w) && isfinite(u.x) &&
              isfinite(u.y) && isfinite(u.z) && isfinite(u.w))) {
            status = 2;
        } else if (!(pp.w > 0.f)) {
            status = 1;
        } else {
            const double q[3] = {pp.x, pp.y, pp.z};
            double F[3] = {0, 0, 0}, H[6] = {0, 0, 0, 0, 0, 0};
            for (int j = 0; j < g.na; ++j) {
                PairQ pq;
                pair_q(q, A[j], M[j], g, pq);
                for (int i = 0; i < 3; ++i) F[i] += pq.F[i];
                for (int i = 0; i < 6; ++i) H[i] += pq.H[i];
            }
            const float4 P0 = pts[ip];
            double tip[3];
            for (int i = 0; i < 3; ++i) tip[i] = (double)(&P0.x)[i] + g.R[2][i] * g.dtip * (double)iz;
            const double dw[3] = {q[0] - tip[0], q[1] - tip[1], q[2] - tip[2]};
            double d[3];
            for (int i = 0; i < 3; ++i) d[i] = g.R[i][0] * dw[0] + g.R[i][1] * dw[1] + g.R[i][2] * dw[2];
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            r = fmax(r, 1e-10);
            const double phi = g.k[3] * (1.0 - g.d0[3] / r), beta = g.k[3] * g.d0[3] / (r * r * r);
            double gt[3], Jt[3][3];
            for (int i = 0; i < 3; ++i) {
                gt[i] = (d[i] - g.d0[i]) * g.k[i] + d[i] * (g.k[3] * (r - g.d0[3]) / r);
                for (int j = 0; j < 3; ++j) Jt[i][j] = beta * d[i] * d[j] + (i == j ? g.k[i] + phi : 0.0);
            }
            const double Hm[3][3] = {{H[0], H[1], H[2]}, {H[1], H[3], H[4]}, {H[2], H[4], H[5]}};
            double G[3], J[3][3];
            for (int i = 0; i < 3; ++i) {
                G[i] = F[i] + g.R[0][i] * gt[0] + g.R[1][i] * gt[1] + g.R[2][i] * gt[2] + g.R[2][i] * g.surf;
                for (int j = 0; j < 3; ++j) {
                    double acc = -Hm[i][j];
                    for (int m = 0; m < 3; ++m)
                        for (int n = 0; n < 3; ++n) acc += g.R[m][i] * Jt[m][n] * g.R[n][j];
                    J[i][j] = acc;
                }
            }
            const double res = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
            dg.x = (float)res;
            bool hfin = true;
            for (int i = 0; i < 6; ++i) hfin = hfin && isfinite(H[i]);
            double uwv[3];
            for (int i = 0; i < 3; ++i) uwv[i] = g.R[0][i] * u.x + g.R[1][i] * u.y + g.R[2][i] * u.z;
            if (!(res < 5e-4)) {
                status = 1;
            } else if (!hfin) {
                status = 2;
            } else {
                double bv[3];
                for (int i = 0; i < 3; ++i)
                    bv[i] = -(Hm[i][0] * uwv[0] + Hm[i][1] * uwv[1] + Hm[i][2] * uwv[2]) - F[i] * (double)u.w;
                double cof[3][3];
                cof[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
                cof[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
                cof[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
                cof[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
                cof[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
                cof[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
                cof[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
                cof[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
                cof[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                const double det = J[0][0] * cof[0][0] + J[0][1] * cof[0][1] + J[0][2] * cof[0][2];
                const double sc = fmax(fmax(fabs(J[0][0]), fabs(J[1][1])), fmax(fabs(J[2][2]), 1e-20));
                double l[3] = {0, 0, 0};
                if (!(fabs(det) >= 1e-18 * sc * sc * sc)) {
                    status = 4;
                } else {
                    // J lambda = b  =>  lambda = adj(J) b / det, adj(J) = cof^T
                    for (int i = 0; i < 3; ++i) l[i] = (cof[0][i] * bv[0] + cof[1][i] * bv[1] + cof[2][i] * bv[2]) / det;
                    if (!(isfinite(l[0]) && isfinite(l[1]) && isfinite(l[2]))) {
                        status = 2;
                    } else {
                        double S[3][3];
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) S[i][j] = -0.5 * (J[i][j] + J[j][i]);
                        const double lmin = eigmin_sym3(S);
                        dg.y = (float)lmin;
                        double jn = 0.0, an = 0.0;
                        for (int i = 0; i < 3; ++i) {
                            jn = fmax(jn, fabs(J[i][0]) + fabs(J[i][1]) + fabs(J[i][2]));
                            an = fmax(an, fabs(cof[0][i]) + fabs(cof[1][i]) + fabs(cof[2][i]));
                        }
                        const double cond = jn * an / fmax(fabs(det), 1e-300);
                        if (!(lmin > 1e-5)) {
                            status = 3;
                        } else {
                            dg.z = (float)cond;
                            if (!(cond < 1e6)) status = 4;
                        }
                    }
                }
                if (status == 0) {
                    lam = make_float4((float)l[0], (float)l[1], (float)l[2], 0.f);
                    me.ok = 1;
                    for (int i = 0; i < 3; ++i) { me.q[i] = q[i]; me.w[i] = uwv[i] - l[i]; }
                    me.uw = (double)u.w;
                }
            }
        }
        dg.w = (float)status;
        lambdas[s] = lam;
        diag[s] = dg;
    }
    SL[threadIdx.x] = me;
    __syncthreads();

    // phase 2: per-atom partials over the block's states
    const int na = g.na;
    const int nparts = kAdjBlock / na;        // na <= 128  =>  nparts >= 2
    const int a = threadIdx.x % na, pi = threadIdx.x / na;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (pi < nparts) {
        const float4 at = A[a], cm = M[a];
        for (int t = pi; t < kAdjBlock; t += nparts) {
            const StateL& st = SL[t];
            if (!st.ok) continue;
            PairQ pq;
            pair_q(st.q, at, cm, g, pq);
            const double* w = st.w;
            acc[0] += pq.H[0] * w[0] + pq.H[1] * w[1] + pq.H[2] * w[2] + pq.F[0] * st.uw;
            acc[1] += pq.H[1] * w[0] + pq.H[3] * w[1] + pq.H[4] * w[2] + pq.F[1] * st.uw;
            acc[2] += pq.H[2] * w[0] + pq.H[4] * w[1] + pq.H[5] * w[2] + pq.F[2] * st.uw;
            const double wd = w[0] * pq.d[0] + w[1] * pq.d[1] + w[2] * pq.d[2];
            acc[3] += wd * pq.dFR0 + st.uw * pq.dER0;
            acc[4] += wd * pq.dFE0 + st.uw * pq.dEE0;
            acc[5] += w[0] * pq.dFQ[0] + w[1] * pq.dFQ[1] + w[2] * pq.dFQ[2] + st.uw * pq.dEQ;
        }
    }
    for (int c = 0; c < 6; ++c) red[threadIdx.x][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < na) {
        double sum[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < nparts; ++q)
            for (int c = 0; c < 6; ++c) sum[c] += red[q * na + threadIdx.x][c];
        double* out = part + ((size_t)blockIdx.x * na + threadIdx.x) * 6;
        for (int c = 0; c < 6; ++c) out[c] = sum[c];
    }
}

__global__ __launch_bounds__(256) void k_direct_reduce(const double* __restrict__ part, int nblk, int na,
                                                       double* __restrict__ grad) {
    __shared__ double red[256][6];
    const int a = blockIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += 256) {
        const double* p = part + ((size_t)b * na + a) * 6;
        for (int c = 0; c < 6; ++c) acc[c] += p[c];
    }
    for (int c = 0; c < 6; ++c) red[threadIdx.x][c] = acc[c];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int c = 0; c < 6; ++c) red[threadIdx.x][c] += red[threadIdx.x + h][c];
        __syncthreads();
    }
    if (threadIdx.x < 6) grad[(size_t)a * 6 + threadIdx.x] = red[0][threadIdx.x];
}

int n_adj_blocks(int n_scan, int nz) { return (int)(((size_t)n_scan * nz + kAdjBlock - 1) / kAdjBlock); }

}  // namespace

size_t direct_vjp_workspace(int na, int n_scan, int nz) {
    return (size_t)n_adj_blocks(n_scan, nz) * (size_t)na * 6 * sizeof(double);
}

int launch_direct_relax(const float* atoms, const float* cms, int na, const float* pts, int n_scan, int nz,
                        const PpmTip& tip, const float Qs[4], const float QZs[4], const PpmRelaxCtl& ctl, float* FEs,
                        float* PPs, hipStream_t stream) {
    DirectArgs d{};
    d.na = na;
    for (int k = 0; k < 4; ++k)
        if (Qs[k] != 0.f) {
            d.Qs[d.nsite] = Qs[k] * kCOULOMB;
            d.QZs[d.nsite] = QZs[k];
            ++d.nsite;
        }
    rc::TiltArgs p{};
    p.n_scan = n_scan; p.nz = nz; p.max_iter = ctl.max_iter; p.check_before = 1;
    p.scan_ny = ctl.scan_ny; p.f2conv = ctl.f2conv; p.damped = ctl.damped != 0;
    for (int q = 0; q < 3; ++q) { p.A[q] = tip.tipA[q]; p.B[q] = tip.tipB[q]; p.C[q] = tip.tipC[q]; }
    p.dtip = tip.tipC[3];
    for (int q = 0; q < 4; ++q) { p.k[q] = tip.stiffness[q]; p.d0[q] = tip.dpos0[q]; }
    p.dt0 = tip.relax[0]; p.damp0 = tip.relax[1]; p.surf = tip.surfFF[0];
    const int g = direct_group(n_scan, na);
    const dim3 blocks((unsigned)(((long long)n_scan * g + 255) / 256));
    const float4* A = reinterpret_cast<const float4*>(atoms);
    const float4* M = reinterpret_cast<const float4*>(cms);
    const float4* S = reinterpret_cast<const float4*>(pts);
    float4* F = reinterpret_cast<float4*>(FEs);
    float4* Q = reinterpret_cast<float4*>(PPs);
    switch (g) {
        case 1: hipLaunchKernelGGL(k_direct_relax<1>, blocks, dim3(256), 0, stream, A, M, d, S, p, F, Q); break;
        case 2: hipLaunchKernelGGL(k_direct_relax<2>, blocks, dim3(256), 0, stream, A, M, d, S, p, F, Q); break;
        case 4: hipLaunchKernelGGL(k_direct_relax<4>, blocks, dim3(256), 0, stream, A, M, d, S, p, F, Q); break;
        case 8: hipLaunchKernelGGL(k_direct_relax<8>, blocks, dim3(256), 0, stream, A, M, d, S, p, F, Q); break;
        case 16: hipLaunchKernelGGL(k_direct_relax<16>, blocks, dim3(256), 0, stream, A, M, d, S, p, F, Q); break;
        case 32: hipLaunchKernelGGL(k_direct_relax<32>, blocks, dim3(256), 0, stream, A, M, d, S, p, F, Q); break;
        default: hipLaunchKernelGGL(k_direct_relax<64>, blocks, dim3(256), 0, stream, A, M, d, S, p, F, Q); break;
    }
    return check_launch("direct_relax");
}

int launch_direct_vjp(const float* atoms, const float* cms, int na, const float* pts, int n_scan, int nz,
                      const PpmTip& tip, const float Qs[4], const float QZs[4], const float* PPs, const float* dL_dFEs,
                      float* lambdas, float* diag, void* workspace, double* grad, hipStream_t stream) {
    AdjArgs g{};
    g.na = na; g.n_scan = n_scan; g.nz = nz;
    for (int k = 0; k < 4; ++k)
        if (Qs[k] != 0.f) {
            g.Qs[g.nsite] = (double)Qs[k] * 14.3996448915;
            g.QZs[g.nsite] = QZs[k];
            ++g.nsite;
        }
    for (int i = 0; i < 3; ++i) { g.R[0][i] = tip.tipA[i]; g.R[1][i] = tip.tipB[i]; g.R[2][i] = tip.tipC[i]; }
    g.dtip = tip.tipC[3];
    for (int q = 0; q < 4; ++q) { g.k[q] = tip.stiffness[q]; g.d0[q] = tip.dpos0[q]; }
    g.surf = tip.surfFF[0];
    const int nblk = n_adj_blocks(n_scan, nz);
    double* part = reinterpret_cast<double*>(workspace);
    hipLaunchKernelGGL(k_direct_adjoint, dim3(nblk), dim3(kAdjBlock), 0, stream,
                       reinterpret_cast<const float4*>(atoms), reinterpret_cast<const float4*>(cms),
                       reinterpret_cast<const float4*>(pts), g, reinterpret_cast<const float4*>(PPs),
                       reinterpret_cast<const float4*>(dL_dFEs), reinterpret_cast<float4*>(lambdas),
                       reinterpret_cast<float4*>(diag), part);
    if (int r = check_launch("direct_adjoint")) return r;
    hipLaunchKernelGGL(k_direct_reduce, dim3(na), dim3(256), 0, stream, part, nblk, na, grad);
    return check_launch("direct_reduce");
}

}  // namespace ppm
