// ppm_ewald.hip — 2D Ewald electrostatics of a periodic slab (spammm/surfaces/Ewald2D.py,
// spammm/surfaces/SurfaceEwald.py, kernels/surface.cl:1036-1381):
//   phi(rho, z) = C [ -(2 pi / A) sum_i q_i |z - z_i|
//                     + sum_{G != 0} sum_i (2 pi / (A |G|)) q_i Re e^{i G.(rho - rho_i)} e^{-|G| |z - z_i|} ]
// G = h b1 + k b2, |h|,|k| <= n_harm.  Ions sharing one height form a layer; per layer l and G
// the structure factor W_l(G) = (2 pi / (A |G|)) sum_{i in l} q_i e^{-i G.rho_i} is computed
// once in FP64, so a slab of L layers costs L terms per G instead of one per ion.  G whose
// structure factors vanish in every layer (a supercell of a smaller lattice keeps only the
// small lattice's reciprocal vectors) are dropped by an ordered compaction, so evaluation runs
// over the kept list only.
//
// Prepared system (one device buffer, ppm_ewald_system_bytes):
//   double2 W[n_layers][n_g] | double zl[n_layers] | double Ql[n_layers] | int cnt[64] |
//   int list[n_g] | float4 ginfo[n_g] = (Gx, Gy, h, k) of the kept G, in kept order
//
// Probe grid: per grid plane z and kept G the tables
//   T_G(z)  = sum_k C Qs_k sum_l W_l(G) e^{-|G| |z + QZs_k - z_l|},  T'_G(z) = dT_G/dz
//   t0(z)   = -C (2 pi / A) sum_k Qs_k sum_l Q_l |z + QZs_k - z_l|,   t0'(z) = dt0/dz
// are built in FP64 and stored in FP32; a voxel then needs its in-plane phase p = e^{i G.rho}
// (one sincospi per G and column, shared by kVz planes) and
//   E = t0 + sum_G Re(T_G p),  F = (sum_G Gx Im(T_G p), sum_G Gy Im(T_G p), -t0' - sum_G Re(T'_G p)),
// i.e. F = -grad E.  A workgroup stages the tables of its planes through LDS in chunks of G.
#include "ppm_common.hpp"

namespace ppm {
namespace {

constexpr double kPi = 3.141592653589793;
constexpr int kGChunk = 64;
constexpr int kVz = 16;          // z planes per thread of the grid kernel
constexpr int kScanThreads = 1024;

struct EwGeom {
    int n_harm, n_g, n_layers;
    double b1[2], b2[2], area;
};

struct SysView {
    double2* W;
    double* zl;
    double* Ql;
    int* cnt;
    int* list;
    float4* ginfo;
};

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

inline SysView sys_view(void* base, int n_layers, int n_g) {
    char* p = static_cast<char*>(base);
    SysView v;
    size_t o = 0;
    v.W = reinterpret_cast<double2*>(p + o);
    o += align256((size_t)n_layers * n_g * sizeof(double2));
    v.zl = reinterpret_cast<double*>(p + o);
    o += align256((size_t)n_layers * sizeof(double));
    v.Ql = reinterpret_cast<double*>(p + o);
    o += align256((size_t)n_layers * sizeof(double));
    v.cnt = reinterpret_cast<int*>(p + o);
    o += 256;
    v.list = reinterpret_cast<int*>(p + o);
    o += align256((size_t)n_g * sizeof(int));
    v.ginfo = reinterpret_cast<float4*>(p + o);
    return v;
}

size_t sys_bytes(int n_layers, int n_g) {
    return align256((size_t)n_layers * n_g * sizeof(double2)) + 2 * align256((size_t)n_layers * sizeof(double)) +
           256 + align256((size_t)n_g * sizeof(int)) + align256((size_t)n_g * sizeof(float4));
}

// G vectors in the order of generate_G_vectors (spammm/surfaces/Ewald2D.py:63-84):
// h = -n..n outer, k = -n..n inner, (0,0) skipped.
__host__ __device__ inline void g_of(int n_harm, int g, int& h, int& k) {
    const int w = 2 * n_harm + 1;
    const int idx = g >= (w * w - 1) / 2 ? g + 1 : g;
    h = idx / w - n_harm;
    k = idx % w - n_harm;
}

__device__ inline void g_vec(const EwGeom& a, int h, int k, double& Gx, double& Gy, double& Gn) {
    Gx = h * a.b1[0] + k * a.b2[0];
    Gy = h * a.b1[1] + k * a.b2[1];
    Gn = sqrt(Gx * Gx + Gy * Gy);
}

// Phase fraction h s1 + k s2 mod 1 (s = fractional coordinates in [0,1)), each product reduced
// before the sum so that large h, k keep full FP64 accuracy.
__device__ inline double phase_frac(int h, int k, double s1, double s2) {
    const double f1 = h * s1, f2 = k * s2;
    return (f1 - floor(f1)) + (f2 - floor(f2));
}

// W[l][g] (compute_w_per_ion summed per layer, spammm/surfaces/Ewald2D.py:103-116).
__global__ __launch_bounds__(256) void k_layer_coeffs(const float4* __restrict__ ions, const int* __restrict__ off,
                                                      EwGeom a, double2* __restrict__ W) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.n_layers * a.n_g) return;
    const int l = t / a.n_g, g = t - l * a.n_g;
    int h, k;
    g_of(a.n_harm, g, h, k);
    double Gx, Gy, Gn;
    g_vec(a, h, k, Gx, Gy, Gn);
    double re = 0.0, im = 0.0;
    for (int i = off[l]; i < off[l + 1]; ++i) {
        const float4 q = ions[i];
        double s1 = (a.b1[0] * (double)q.x + a.b1[1] * (double)q.y) * (0.5 / kPi);
        double s2 = (a.b2[0] * (double)q.x + a.b2[1] * (double)q.y) * (0.5 / kPi);
        s1 -= floor(s1);
        s2 -= floor(s2);
        double sn, cs;
        sincospi(2.0 * phase_frac(h, k, s1, s2), &sn, &cs);
        re += (double)q.w * cs;
        im -= (double)q.w * sn;
    }
    const double pref = 2.0 * kPi / (a.area * Gn);
    W[t] = make_double2(pref * re, pref * im);
}

// Layer heights and total charges.
__global__ void k_layer_sums(const float4* __restrict__ ions, const int* __restrict__ off, int n_layers,
                             double* __restrict__ zl, double* __restrict__ Ql) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_layers) return;
    double q = 0.0;
    for (int i = off[l]; i < off[l + 1]; ++i) q += (double)ions[i].w;
    zl[l] = (double)ions[off[l]].z;
    Ql[l] = q;
}

__device__ inline double g_weight2(const double2* __restrict__ W, int n_layers, int n_g, int g) {
    double mg = 0.0;
    for (int l = 0; l < n_layers; ++l) {
        const double2 w = W[(size_t)l * n_g + g];
        mg = fmax(mg, w.x * w.x + w.y * w.y);
    }
    return mg;
}

// Ordered compaction (one workgroup): keep G with max_l |W_l(G)| > drop_tol * max_{l,G} |W_l(G)|.
// G and -G contribute equal terms: the charges are real, so W_l(-G) = conj(W_l(G)) and
// Re(W_l(-G) e^{-iG.r}) = Re(W_l(G) e^{iG.r}), with the same -grad.  The kept list holds the half
// plane h > 0 or (h = 0, k > 0) of the G above the drop threshold (|W(-G)| = |W(G)|, so pairs are
// kept or dropped together) and every consumer counts each entry twice.
__device__ inline bool kept_g(const double2* __restrict__ W, const EwGeom& a, int g, double thr) {
    int h, k;
    g_of(a.n_harm, g, h, k);
    if (!(h > 0 || (h == 0 && k > 0))) return false;
    const double mg = g_weight2(W, a.n_layers, a.n_g, g);
    return mg > thr && mg > 0.0;
}

__global__ __launch_bounds__(kScanThreads) void k_compact(const double2* __restrict__ W, EwGeom a, double drop_tol,
                                                          int* __restrict__ cnt, int* __restrict__ list,
                                                          float4* __restrict__ ginfo) {
    __shared__ double red[kScanThreads / 64];
    __shared__ int scan[kScanThreads];
    const int tid = threadIdx.x;
    const long long nw = (long long)a.n_layers * a.n_g;
    double m = 0.0;
    for (long long i = tid; i < nw; i += kScanThreads) {
        const double2 w = W[i];
        m = fmax(m, w.x * w.x + w.y * w.y);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    double wmax = 0.0;
    for (int i = 0; i < kScanThreads / 64; ++i) wmax = fmax(wmax, red[i]);
    const double thr = drop_tol * drop_tol * wmax;     // compared on |W|^2
    const int per = (a.n_g + kScanThreads - 1) / kScanThreads;
    const int g0 = min(a.n_g, tid * per), g1 = min(a.n_g, g0 + per);
    int mine = 0;
    for (int g = g0; g < g1; ++g) mine += kept_g(W, a, g, thr) ? 1 : 0;
    scan[tid] = mine;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {       // inclusive Hillis-Steele scan
        const int v = tid >= o ? scan[tid - o] : 0;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    int pos = scan[tid] - mine;
    for (int g = g0; g < g1; ++g) {
        if (kept_g(W, a, g, thr)) {
            int h, k;
            g_of(a.n_harm, g, h, k);
            double Gx, Gy, Gn;
            g_vec(a, h, k, Gx, Gy, Gn);
            list[pos] = g;
            ginfo[pos] = make_float4((float)Gx, (float)Gy, (float)h, (float)k);
            ++pos;
        }
    }
    if (tid == kScanThreads - 1) cnt[0] = scan[tid];
}

// Plane tables: tab[iz][c] = (T.re, T.im, T'.re, T'.im) for kept G c < cnt; zt[iz] = (t0, t0').
__global__ __launch_bounds__(256) void k_plane_tables(SysView s, EwGeom a, int nz, double z0, double dz,
                                                      double4 qs, double4 qz, float4* __restrict__ tab,
                                                      float2* __restrict__ zt) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nz * (a.n_g + 1)) return;
    const int iz = (int)(t / (a.n_g + 1)), c = (int)(t - (long long)iz * (a.n_g + 1));
    const double z = z0 + dz * iz;
    const double qv[4] = {qs.x, qs.y, qs.z, qs.w}, zv[4] = {qz.x, qz.y, qz.z, qz.w};
    if (c == a.n_g) {                                   // the G = 0 term
        double t0 = 0.0, t1 = 0.0;
        for (int q = 0; q < 4; ++q)
            for (int l = 0; l < a.n_layers; ++l) {
                const double d = z + zv[q] - s.zl[l];
                t0 += qv[q] * s.Ql[l] * fabs(d);
                t1 += qv[q] * s.Ql[l] * (d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0));
            }
        const double f = -2.0 * kPi / a.area * kCOULOMB_D;
        zt[iz] = make_float2((float)(f * t0), (float)(f * t1));
        return;
    }
    if (c >= min(s.cnt[0], a.n_g)) return;
    const int g = min(max(s.list[c], 0), a.n_g - 1);
    int h, k;
    g_of(a.n_harm, g, h, k);
    double Gx, Gy, Gn;
    g_vec(a, h, k, Gx, Gy, Gn);
    double tr = 0.0, ti = 0.0, dr = 0.0, di = 0.0;
    for (int q = 0; q < 4; ++q) {
        if (qv[q] == 0.0) continue;
        for (int l = 0; l < a.n_layers; ++l) {
            const double d = z + zv[q] - s.zl[l];
            const double e = qv[q] * kCOULOMB_D * exp(-Gn * fabs(d));
            const double2 w = s.W[(size_t)l * a.n_g + g];
            const double sg = d > 0.0 ? -Gn : (d < 0.0 ? Gn : 0.0);
            tr += e * w.x;
            ti += e * w.y;
            dr += sg * e * w.x;
            di += sg * e * w.y;
        }
    }
    // x2: the entry stands for the pair +-G
    tab[(size_t)iz * a.n_g + c] = make_float4((float)(2.0 * tr), (float)(2.0 * ti), (float)(2.0 * dr),
                                              (float)(2.0 * di));
}

// One thread = one (ia, ib) column x kVz planes.  Voxel position p0 + ia dA + ib dB + ic dC with
// dA, dB in the xy plane and dC along z.  out[(ic*nb + ib)*na + ia] (+)= (Fx, Fy, Fz, E).
// Every thread of a workgroup takes part in the LDS staging; out-of-range columns only skip the
// store.
__global__ __launch_bounds__(256) void k_ewald_grid(const float4* __restrict__ tab, const float2* __restrict__ zt,
                                                    const float4* __restrict__ ginfo, const int* __restrict__ cntp,
                                                    int n_g, int na, int nb, int nz, float px0, float py0,
                                                    float dax, float day, float dbx, float dby, float b1x, float b1y,
                                                    float b2x, float b2y, int accumulate, float4* __restrict__ out) {
    __shared__ float4 st[kVz][kGChunk];
    __shared__ float4 sg[kGChunk];
    const int col = blockIdx.x * 256 + threadIdx.x;
    const int iz0 = blockIdx.y * kVz;
    const bool ok = col < na * nb;
    const int cc = ok ? col : 0;
    const int ia = cc % na, ib = cc / na;
    const float x = px0 + dax * (float)ia + dbx * (float)ib;
    const float y = py0 + day * (float)ia + dby * (float)ib;
    const float inv2pi = 0.15915494309189535f;
    float s1 = (b1x * x + b1y * y) * inv2pi;          // fractional coordinates along a, b
    float s2 = (b2x * x + b2y * y) * inv2pi;
    s1 -= floorf(s1);
    s2 -= floorf(s2);
    const int cnt = min(max(cntp[0], 0), n_g);
    float e[kVz], fx[kVz], fy[kVz], fz[kVz];
#pragma unroll
    for (int v = 0; v < kVz; ++v) { e[v] = 0.f; fx[v] = 0.f; fy[v] = 0.f; fz[v] = 0.f; }
    for (int c0 = 0; c0 < cnt; c0 += kGChunk) {
        const int m = min(kGChunk, cnt - c0);
        __syncthreads();
        for (int u = threadIdx.x; u < kVz * kGChunk; u += 256) {
            const int v = u / kGChunk, j = u - v * kGChunk;
            const int iz = min(iz0 + v, nz - 1);
            st[v][j] = j < m ? tab[(size_t)iz * n_g + c0 + j] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (threadIdx.x < kGChunk)
            sg[threadIdx.x] = (int)threadIdx.x < m ? ginfo[c0 + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        for (int j = 0; j < m; ++j) {
            const float4 G = sg[j];
            float f1 = G.z * s1, f2 = G.w * s2;
            const float f = (f1 - rintf(f1)) + (f2 - rintf(f2));
            float ps, pc;
            sincospif(2.f * f, &ps, &pc);
#pragma unroll
            for (int v = 0; v < kVz; ++v) {
                const float4 T = st[v][j];
                e[v] = fmaf(T.x, pc, fmaf(-T.y, ps, e[v]));
                const float im = fmaf(T.x, ps, T.y * pc);
                fx[v] = fmaf(G.x, im, fx[v]);
                fy[v] = fmaf(G.y, im, fy[v]);
                fz[v] = fmaf(-T.z, pc, fmaf(T.w, ps, fz[v]));
            }
        }
    }
    if (!ok) return;
#pragma unroll
    for (int v = 0; v < kVz; ++v) {
        const int iz = iz0 + v;
        if (iz < nz) {
            const float2 t0 = zt[iz];
            float4 r = make_float4(fx[v], fy[v], fz[v] - t0.y, e[v] + t0.x);
            const size_t o = ((size_t)iz * nb + ib) * na + ia;
            if (accumulate) {
                const float4 b = out[o];
                r.x += b.x; r.y += b.y; r.z += b.z; r.w += b.w;
            }
            out[o] = r;
        }
    }
}

// phi at points (n,4) = (x,y,z,-), C included, FP64 accumulation over the kept G.
// mode 0: the full form (eval_potential_full, kernels/surface.cl:1172-1246), any z;
// mode 1: the vacuum form (eval_potential_vacuum, kernels/surface.cl:1103-1157): no G = 0 term,
// decay e^{-|G| (z - z_l)} (z above every layer).
__global__ __launch_bounds__(256) void k_ewald_points(const float4* __restrict__ P, int np, SysView s, EwGeom a,
                                                      int mode, float* __restrict__ phi) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const float4 p = P[i];
    const double z = (double)p.z;
    double s1 = (a.b1[0] * (double)p.x + a.b1[1] * (double)p.y) * (0.5 / kPi);
    double s2 = (a.b2[0] * (double)p.x + a.b2[1] * (double)p.y) * (0.5 / kPi);
    s1 -= floor(s1);
    s2 -= floor(s2);
    double acc = 0.0;
    if (mode == 0)
        for (int l = 0; l < a.n_layers; ++l) acc -= (2.0 * kPi / a.area) * s.Ql[l] * fabs(z - s.zl[l]);
    const int cnt = min(s.cnt[0], a.n_g);
    for (int c = 0; c < cnt; ++c) {
        const int g = min(max(s.list[c], 0), a.n_g - 1);
        int h, k;
        g_of(a.n_harm, g, h, k);
        double Gx, Gy, Gn;
        g_vec(a, h, k, Gx, Gy, Gn);
        double sn, cs;
        sincospi(2.0 * phase_frac(h, k, s1, s2), &sn, &cs);
        for (int l = 0; l < a.n_layers; ++l) {
            const double2 w = s.W[(size_t)l * a.n_g + g];
            const double d = mode == 0 ? fabs(z - s.zl[l]) : z - s.zl[l];
            acc += 2.0 * (w.x * cs - w.y * sn) * exp(-Gn * d);     // the pair +-G
        }
    }
    phi[i] = (float)(acc * kCOULOMB_D);
}

// Direct Coulomb sum over the images n a + m b with n^2 + m^2 <= n_rep^2 (eval_potential_brute,
// kernels/surface.cl:1260-1305) or, square, every |n|,|m| <= n_rep (eval_potential_brute,
// spammm/surfaces/Ewald2D.py:176-204), FP64 accumulation.
__global__ __launch_bounds__(256) void k_ewald_brute(const float4* __restrict__ P, int np,
                                                     const float4* __restrict__ ions, int ni, double ax, double ay,
                                                     double bx, double by, int n_rep, int square,
                                                     float* __restrict__ phi) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const float4 p = P[i];
    double acc = 0.0;
    for (int n = -n_rep; n <= n_rep; ++n)
        for (int m = -n_rep; m <= n_rep; ++m) {
            if (!square && n * n + m * m > n_rep * n_rep) continue;
            const double Rx = n * ax + m * bx, Ry = n * ay + m * by;
            for (int j = 0; j < ni; ++j) {
                const float4 q = ions[j];
                const double dx = (double)p.x - ((double)q.x + Rx), dy = (double)p.y - ((double)q.y + Ry);
                const double dz = (double)p.z - (double)q.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 1e-24) acc += (double)q.w / sqrt(r2);
            }
        }
    phi[i] = (float)(acc * kCOULOMB_D);
}

// Finite-cluster Coulomb sum, no images (eval_potential_cluster, kernels/surface.cl:1321-1381):
// ions staged through LDS, FP64 accumulation.  Every thread of the workgroup takes part in the
// staging; out-of-range threads only skip the store.
__global__ __launch_bounds__(256) void k_ewald_cluster(const float4* __restrict__ P, int np,
                                                       const float4* __restrict__ ions, int ni,
                                                       float* __restrict__ phi) {
    __shared__ float4 sion[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const float4 p = P[i < np ? i : 0];
    double acc = 0.0;
    for (int j0 = 0; j0 < ni; j0 += 256) {
        __syncthreads();
        const int j = j0 + threadIdx.x;
        sion[threadIdx.x] = j < ni ? ions[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const int m = min(256, ni - j0);
        for (int u = 0; u < m; ++u) {
            const float4 q = sion[u];
            const double dx = (double)p.x - q.x, dy = (double)p.y - q.y, dz = (double)p.z - q.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 1e-24) acc += (double)q.w / sqrt(r2);
        }
    }
    if (i < np) phi[i] = (float)(acc * kCOULOMB_D);
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

// Reciprocal lattice of make_reciprocal_2d (spammm/surfaces/Ewald2D.py:39-60).
EwGeom make_geom(int n_harm, int n_layers, const double a[2], const double b[2]) {
    EwGeom g{};
    g.n_harm = n_harm;
    g.n_g = ewald_num_g(n_harm);
    g.n_layers = n_layers;
    g.area = fabs(a[0] * b[1] - a[1] * b[0]);
    g.b1[0] = 2.0 * kPi / g.area * b[1];
    g.b1[1] = -2.0 * kPi / g.area * b[0];
    g.b2[0] = -2.0 * kPi / g.area * a[1];
    g.b2[1] = 2.0 * kPi / g.area * a[0];
    return g;
}

}  // namespace

int ewald_num_g(int n_harm) { return (2 * n_harm + 1) * (2 * n_harm + 1) - 1; }

size_t ewald_system_bytes(int n_layers, int n_harm) { return sys_bytes(n_layers, ewald_num_g(n_harm)); }

size_t ewald_kept_count_offset(int n_layers, int n_harm) {
    const SysView v = sys_view(nullptr, n_layers, ewald_num_g(n_harm));
    return (size_t)(reinterpret_cast<const char*>(v.cnt) - static_cast<const char*>(nullptr));
}

size_t ewald_grid_workspace(int n_harm, int nz) {
    return align256((size_t)nz * ewald_num_g(n_harm) * sizeof(float4)) + align256((size_t)nz * sizeof(float2));
}

int launch_ewald_prepare(const float* ions, const int* layer_off, int n_layers, const double a[2], const double b[2],
                         int n_harm, double drop_tol, void* system, hipStream_t stream) {
    const EwGeom g = make_geom(n_harm, n_layers, a, b);
    const SysView s = sys_view(system, n_layers, g.n_g);
    const float4* ion4 = reinterpret_cast<const float4*>(ions);
    hipLaunchKernelGGL(k_layer_coeffs, dim3(blocks_for((long long)n_layers * g.n_g)), dim3(256), 0, stream, ion4,
                       layer_off, g, s.W);
    if (int r = check_launch("ewald_coeffs")) return r;
    hipLaunchKernelGGL(k_layer_sums, dim3((n_layers + 63) / 64), dim3(64), 0, stream, ion4, layer_off, n_layers, s.zl,
                       s.Ql);
    if (int r = check_launch("ewald_layer_sums")) return r;
    hipLaunchKernelGGL(k_compact, dim3(1), dim3(kScanThreads), 0, stream, s.W, g, drop_tol, s.cnt, s.list, s.ginfo);
    return check_launch("ewald_compact");
}

int launch_ewald_points(int mode, const float* pts, int np, const void* system, int n_layers, int n_harm,
                        const double a[2], const double b[2], float* phi, hipStream_t stream) {
    if (np == 0) return 0;
    const EwGeom g = make_geom(n_harm, n_layers, a, b);
    const SysView s = sys_view(const_cast<void*>(system), n_layers, g.n_g);
    hipLaunchKernelGGL(k_ewald_points, dim3(blocks_for(np)), dim3(256), 0, stream,
                       reinterpret_cast<const float4*>(pts), np, s, g, mode, phi);
    return check_launch("ewald_points");
}

int launch_ewald_brute(const float* pts, int np, const float* ions, int ni, const double a[2], const double b[2],
                       int n_rep, bool square, float* phi, hipStream_t stream) {
    if (np == 0) return 0;
    hipLaunchKernelGGL(k_ewald_brute, dim3(blocks_for(np)), dim3(256), 0, stream,
                       reinterpret_cast<const float4*>(pts), np, reinterpret_cast<const float4*>(ions), ni, a[0], a[1],
                       b[0], b[1], n_rep, square ? 1 : 0, phi);
    return check_launch("ewald_brute");
}

int launch_ewald_cluster(const float* pts, int np, const float* ions, int ni, float* phi, hipStream_t stream) {
    if (np == 0) return 0;
    hipLaunchKernelGGL(k_ewald_cluster, dim3(blocks_for(np)), dim3(256), 0, stream,
                       reinterpret_cast<const float4*>(pts), np, reinterpret_cast<const float4*>(ions), ni, phi);
    return check_launch("ewald_cluster");
}

int launch_ewald_tip_grid(const void* system, int n_layers, int n_harm, const double a[2], const double b[2],
                          const PpmGridSpec& gs, const float Qs[4], const float QZs[4], int accumulate,
                          void* workspace, float* grid, hipStream_t stream) {
    const EwGeom g = make_geom(n_harm, n_layers, a, b);
    const SysView s = sys_view(const_cast<void*>(system), n_layers, g.n_g);
    float4* tab = static_cast<float4*>(workspace);
    float2* zt = reinterpret_cast<float2*>(static_cast<char*>(workspace) +
                                           align256((size_t)gs.nz * g.n_g * sizeof(float4)));
    const double4 qs = make_double4(Qs[0], Qs[1], Qs[2], Qs[3]);
    const double4 qz = make_double4(QZs[0], QZs[1], QZs[2], QZs[3]);
    hipLaunchKernelGGL(k_plane_tables, dim3(blocks_for((long long)gs.nz * (g.n_g + 1))), dim3(256), 0, stream, s, g,
                       gs.nz, (double)gs.p0[2], (double)gs.dC[2], qs, qz, tab, zt);
    if (int r = check_launch("ewald_tables")) return r;
    const dim3 grd(blocks_for((long long)gs.nx * gs.ny), (unsigned)((gs.nz + kVz - 1) / kVz));
    hipLaunchKernelGGL(k_ewald_grid, grd, dim3(256), 0, stream, tab, zt, s.ginfo, s.cnt, g.n_g, gs.nx, gs.ny, gs.nz,
                       gs.p0[0], gs.p0[1], gs.dA[0], gs.dA[1], gs.dB[0], gs.dB[1], (float)g.b1[0], (float)g.b1[1],
                       (float)g.b2[0], (float)g.b2[1], accumulate, reinterpret_cast<float4*>(grid));
    return check_launch("ewald_grid");
}

}  // namespace ppm
