// Parallel spin-current response to an electric field (eintshe,
// eintshe.F90:2-224): per (k, E) pair the Green functions Gq = G(k + q, z + w)
// and G = G(k, z) come from the Green function kernels' raw output, the
// velocity vertex T = dH/dk . u (dtdksub.F90:14-26) is assembled from the hop
// table, and the 16 spin-block terms of the reference are evaluated as
//     df[(s, i, mu, nu), s'] = sum_{(j,g),(l,x)} P[(i,rn,nu),(j,cg,g)] T[(j,g),(l,x)]
//                                                 Q[(l,rx,x),(i,cm,mu)]
// with s -> (rn, cm), s' -> (cg, rx) in {(d,u), (u,u), (d,d), (u,d)}:
// imaginary axis P = Gq, Q = G plus P = Gq^H, Q = G^H (the conjugated second
// product of the reference); real axis P = -i (Gq - Gq^H), Q = G^H.  Per pair
// the product P (1 (x) T) is formed once (d^3 / 2 complex MACs) and only the
// diagonal atom blocks of its product with Q are taken.  Sums over pairs run
// per chunk of pairs (np / 1024 per workgroup), then over chunks, then over
// batches, in a fixed order;
// the RPA prefactor (a D x D matrix, constant over pairs) is applied once to
// the sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "th_internal.h"

namespace {

constexpr int TH_CU_NT = 256;      // threads per workgroup
constexpr int TH_CU_RT = 4;        // 2 x 2 output tiles per thread: D x 4 <= 4 x 4 x 256 for d <= 54
constexpr int TH_CU_WG = 1024;     // workgroups per launch (4 per CU): pairs per workgroup = np / 1024

__device__ inline double2 zmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline void zfma(double2 &s, double2 a, double2 b)
{
    s.x = fma(a.x, b.x, fma(-a.y, b.y, s.x));
    s.y = fma(a.x, b.y, fma(a.y, b.x, s.y));
}
// P~[r][col]: the imaginary axis' P = Gq (term 0) or Gq^H (term 1), the real
// axis' P = -i (Gq - Gq^H); Gq row-major d x d
__device__ inline double2 ptil(const double2 *P, int d, int r, int col, int real, int term)
{
    if (real) {
        const double2 a = P[r * d + col], b = P[col * d + r];
        return make_double2(a.y + b.y, -(a.x - b.x));
    }
    return term == 0 ? P[r * d + col] : zconj(P[col * d + r]);
}
// output tile tt -> atom i, channel s', tile row ta, column tb (2 x 2 tiles of
// the 2 nO x 2 nO block of atom i)
__device__ inline void tile_of(int tt, int nO, int &i, int &sp, int &ta, int &tb)
{
    const int per = 4 * nO * nO;
    i = tt / per;
    const int rem = tt - i * per;
    sp = rem / (nO * nO);
    const int t2 = rem - sp * nO * nO;
    ta = t2 / nO;
    tb = t2 - ta * nO;
}

struct current_params {
    int nA, nO, d, nhop;
    long np, p0, nk_re;
    long chunk;                      // pairs per workgroup
    int real;                        // real-axis pairs (node-major: p = ie * nk_re + ik)
    double u0, u1, u2;               // field direction
    const double *R;
    const double2 *t;                // hops, column-major nO x nO blocks
    const int *hoff;                 // hops of atom pair (j, l): hoff[j nA + l] ..
    const double *kim, *wim, *kre, *wre, *p2;
};

// partial[chunk][o], o = X + D s' (X = 4 ((i nO + mu) nO + nu) + s)
__global__ __launch_bounds__(TH_CU_NT) void current_pair_kernel(current_params c, const double2 *__restrict__ rawq,
                                                                const double2 *__restrict__ raw,
                                                                double2 *__restrict__ partial)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int d = c.d, n = d / 2, nO = c.nO, nA = c.nA;
    double2 *P = reinterpret_cast<double2 *>(smem);   // Gq, d x d row-major
    double2 *Q = P + d * d;                           // G
    double2 *Z = Q + d * d;                           // P~ (1 (x) T): [row][cg n + (l nO + x)]
    double2 *T = Z + d * d;                           // n x n
    double2 *ph = T + n * n;                          // i (u.R) e^{i k.R} per hop
    const int tid = threadIdx.x;
    const int D = 4 * nA * nO * nO, nout = 4 * D, ntile = nout / 4;
    const long b0 = (long)blockIdx.x * c.chunk;
    const long b1 = b0 + c.chunk < c.np ? b0 + c.chunk : c.np;
    double2 acc[4 * TH_CU_RT];
#pragma unroll
    for (int r = 0; r < 4 * TH_CU_RT; ++r) acc[r] = make_double2(0.0, 0.0);
    for (long pl = b0; pl < b1; ++pl) {
        const long p = c.p0 + pl;
        double k0, k1, k2, w;
        if (c.real) {
            const long ie = p / c.nk_re, ik = p - (p / c.nk_re) * c.nk_re;
            k0 = c.kre[3 * ik]; k1 = c.kre[3 * ik + 1]; k2 = c.kre[3 * ik + 2];
            w = c.p2[ie] * c.wre[ik];
        } else {
            k0 = c.kim[3 * p]; k1 = c.kim[3 * p + 1]; k2 = c.kim[3 * p + 2];
            w = c.wim[p];
        }
        const double2 *gq = rawq + (size_t)pl * d * d, *gg = raw + (size_t)pl * d * d;
        for (int e = tid; e < d * d; e += TH_CU_NT) {
            P[e] = gq[e];
            Q[e] = gg[e];
        }
        for (int m = tid; m < c.nhop; m += TH_CU_NT) {
            const double a = k0 * c.R[3 * m] + k1 * c.R[3 * m + 1] + k2 * c.R[3 * m + 2];
            const double v = c.u0 * c.R[3 * m] + c.u1 * c.R[3 * m + 1] + c.u2 * c.R[3 * m + 2];
            double sn, cs;
            sincos(a, &sn, &cs);
            ph[m] = make_double2(-v * sn, v * cs);   // i v (cos a + i sin a)
        }
        __syncthreads();
        for (int e = tid; e < n * n; e += TH_CU_NT) {
            const int rg = e / n, cx = e - (e / n) * n;
            const int j = rg / nO, g = rg - j * nO, l = cx / nO, x = cx - l * nO;
            double2 s = make_double2(0.0, 0.0);
            for (int m = c.hoff[j * nA + l]; m < c.hoff[j * nA + l + 1]; ++m) {
                const double2 tv = zmul(c.t[(size_t)m * nO * nO + g + nO * x], ph[m]);
                s.x += tv.x;
                s.y += tv.y;
            }
            T[e] = s;
        }
        __syncthreads();
        const int nterm = c.real ? 1 : 2;
        for (int term = 0; term < nterm; ++term) {
            // Z[r][cg n + lx] = sum_{jg} P~[r][(2 j + cg) nO + g] T[jg][lx], in 2 x 2 tiles
            for (int tz = tid; tz < (d / 2) * (d / 2); tz += TH_CU_NT) {
                const int r0 = 2 * (tz / (d / 2)), cA = 2 * (tz - (tz / (d / 2)) * (d / 2)), cB = cA + 1;
                const int gA = cA / n, lA = cA - gA * n, gB = cB / n, lB = cB - gB * n;
                double2 z00 = make_double2(0.0, 0.0), z01 = z00, z10 = z00, z11 = z00;
                for (int jg = 0; jg < n; ++jg) {
                    const int j = jg / nO, g = jg - j * nO;
                    const int colA = (2 * j + gA) * nO + g, colB = (2 * j + gB) * nO + g;
                    const double2 p0A = ptil(P, d, r0, colA, c.real, term), p1A = ptil(P, d, r0 + 1, colA, c.real, term);
                    const double2 p0B = gB == gA ? p0A : ptil(P, d, r0, colB, c.real, term);
                    const double2 p1B = gB == gA ? p1A : ptil(P, d, r0 + 1, colB, c.real, term);
                    const double2 tA = T[jg * n + lA], tB = T[jg * n + lB];
                    zfma(z00, p0A, tA);
                    zfma(z01, p0B, tB);
                    zfma(z10, p1A, tA);
                    zfma(z11, p1B, tB);
                }
                Z[r0 * d + cA] = z00;
                Z[r0 * d + cB] = z01;
                Z[(r0 + 1) * d + cA] = z10;
                Z[(r0 + 1) * d + cB] = z11;
            }
            __syncthreads();
            const bool herm = c.real || term == 1;   // Q~ = G^H
            // per atom i and channel s' = (cg, rx): O[a][b] = sum_lx Z[2 i nO + a][cg n + lx]
            // Q~[(2 l + rx) nO + x][2 i nO + b], a = (rn, nu), b = (cm, mu), in 2 x 2 tiles
#pragma unroll
            for (int rr = 0; rr < TH_CU_RT; ++rr) {
                const int tt = tid + TH_CU_NT * rr;
                if (tt < ntile) {
                    int i, sp, ta, tb;
                    tile_of(tt, nO, i, sp, ta, tb);
                    const int cg = 1 - (sp & 1), rx = sp >> 1;
                    const int a0 = 2 * i * nO + 2 * ta, b0 = 2 * i * nO + 2 * tb;
                    const double2 *z0 = Z + (size_t)a0 * d + cg * n, *z1 = z0 + d;
                    double2 o00 = make_double2(0.0, 0.0), o01 = o00, o10 = o00, o11 = o00;
                    for (int lx = 0; lx < n; ++lx) {
                        const int l = lx / nO, x = lx - l * nO, qr = (2 * l + rx) * nO + x;
                        const double2 q0 = herm ? zconj(Q[b0 * d + qr]) : Q[qr * d + b0];
                        const double2 q1 = herm ? zconj(Q[(b0 + 1) * d + qr]) : Q[qr * d + b0 + 1];
                        const double2 za = z0[lx], zb = z1[lx];
                        zfma(o00, za, q0);
                        zfma(o01, za, q1);
                        zfma(o10, zb, q0);
                        zfma(o11, zb, q1);
                    }
                    acc[4 * rr].x += w * o00.x; acc[4 * rr].y += w * o00.y;
                    acc[4 * rr + 1].x += w * o01.x; acc[4 * rr + 1].y += w * o01.y;
                    acc[4 * rr + 2].x += w * o10.x; acc[4 * rr + 2].y += w * o10.y;
                    acc[4 * rr + 3].x += w * o11.x; acc[4 * rr + 3].y += w * o11.y;
                }
            }
            __syncthreads();   // Z is rewritten by the next term or pair
        }
    }
    // tile entry (a, b) -> X = 4 ((i nO + mu) nO + nu) + s, s = 2 cm + 1 - rn; o = X + D s'
#pragma unroll
    for (int rr = 0; rr < TH_CU_RT; ++rr) {
        const int tt = tid + TH_CU_NT * rr;
        if (tt < ntile) {
            int i, sp, ta, tb;
            tile_of(tt, nO, i, sp, ta, tb);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int a = 2 * ta + (e >> 1), b = 2 * tb + (e & 1);
                const int rn = a / nO, nu = a - rn * nO, cm = b / nO, mu = b - cm * nO;
                const int X = 4 * ((i * nO + mu) * nO + nu) + 2 * cm + 1 - rn;
                partial[(size_t)blockIdx.x * nout + X + (size_t)D * sp] = acc[4 * rr + e];
            }
        }
    }
}

// Any d (the kernel above keeps P, Q and Z in LDS and the D x 4 outputs in
// registers, which holds to d = 54): the same sums in three stages through
// device memory.  T per pair, then Z = P~ (1 (x) T) per pair and term, then
// per chunk of pairs the diagonal atom blocks of Z Q~, with the same tiles,
// sums and order of pairs as current_pair_kernel.

// T[pl][jg][lx] = sum over hops (j, l) of t i (u.R) e^{i k.R}
__global__ __launch_bounds__(256) void current_t_kernel(current_params c, double2 *__restrict__ Tg)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double2 *ph = reinterpret_cast<double2 *>(smem);
    const int n = c.d / 2, nO = c.nO, nA = c.nA, tid = threadIdx.x;
    const long pl = blockIdx.x, p = c.p0 + pl;
    double k0, k1, k2;
    if (c.real) {
        const long ik = p - (p / c.nk_re) * c.nk_re;
        k0 = c.kre[3 * ik]; k1 = c.kre[3 * ik + 1]; k2 = c.kre[3 * ik + 2];
    } else {
        k0 = c.kim[3 * p]; k1 = c.kim[3 * p + 1]; k2 = c.kim[3 * p + 2];
    }
    for (int m = tid; m < c.nhop; m += 256) {
        const double a = k0 * c.R[3 * m] + k1 * c.R[3 * m + 1] + k2 * c.R[3 * m + 2];
        const double v = c.u0 * c.R[3 * m] + c.u1 * c.R[3 * m + 1] + c.u2 * c.R[3 * m + 2];
        double sn, cs;
        sincos(a, &sn, &cs);
        ph[m] = make_double2(-v * sn, v * cs);
    }
    __syncthreads();
    double2 *T = Tg + (size_t)pl * n * n;
    for (int e = tid; e < n * n; e += 256) {
        const int rg = e / n, cx = e - (e / n) * n;
        const int j = rg / nO, g = rg - j * nO, l = cx / nO, x = cx - l * nO;
        double2 s = make_double2(0.0, 0.0);
        for (int m = c.hoff[j * nA + l]; m < c.hoff[j * nA + l + 1]; ++m) {
            const double2 tv = zmul(c.t[(size_t)m * nO * nO + g + nO * x], ph[m]);
            s.x += tv.x;
            s.y += tv.y;
        }
        T[e] = s;
    }
}

// Z[pl][r][cg n + lx] = sum_{jg} P~[r][(2 j + cg) nO + g] T[jg][lx], 2 x 2 tiles
// (blockIdx.x = pair, blockIdx.y = group of 256 tiles)
__global__ __launch_bounds__(256) void current_z_kernel(current_params c, const double2 *__restrict__ rawq,
                                                        const double2 *__restrict__ Tg, int term,
                                                        double2 *__restrict__ Zg)
{
    const int d = c.d, n = d / 2, nO = c.nO, h2 = d / 2;
    const long pl = blockIdx.x;
    const int tz = blockIdx.y * 256 + threadIdx.x;
    if (tz >= h2 * h2) return;
    const double2 *P = rawq + (size_t)pl * d * d, *T = Tg + (size_t)pl * n * n;
    double2 *Z = Zg + (size_t)pl * d * d;
    const int r0 = 2 * (tz / h2), cA = 2 * (tz - (tz / h2) * h2), cB = cA + 1;
    const int gA = cA / n, lA = cA - gA * n, gB = cB / n, lB = cB - gB * n;
    double2 z00 = make_double2(0.0, 0.0), z01 = z00, z10 = z00, z11 = z00;
    for (int jg = 0; jg < n; ++jg) {
        const int j = jg / nO, g = jg - j * nO;
        const int colA = (2 * j + gA) * nO + g, colB = (2 * j + gB) * nO + g;
        const double2 p0A = ptil(P, d, r0, colA, c.real, term), p1A = ptil(P, d, r0 + 1, colA, c.real, term);
        const double2 p0B = gB == gA ? p0A : ptil(P, d, r0, colB, c.real, term);
        const double2 p1B = gB == gA ? p1A : ptil(P, d, r0 + 1, colB, c.real, term);
        const double2 tA = T[jg * n + lA], tB = T[jg * n + lB];
        zfma(z00, p0A, tA);
        zfma(z01, p0B, tB);
        zfma(z10, p1A, tA);
        zfma(z11, p1B, tB);
    }
    Z[r0 * d + cA] = z00;
    Z[r0 * d + cB] = z01;
    Z[(r0 + 1) * d + cA] = z10;
    Z[(r0 + 1) * d + cB] = z11;
}

// partial[chunk][o] (+)= sum over the chunk's pairs of w O, O the diagonal
// atom blocks of Z Q~ in 2 x 2 tiles (one tile per thread, blockIdx.y = chunk)
__global__ __launch_bounds__(256) void current_o_kernel(current_params c, const double2 *__restrict__ Zg,
                                                        const double2 *__restrict__ raw, int herm, int add,
                                                        double2 *__restrict__ partial)
{
    const int d = c.d, n = d / 2, nO = c.nO, nA = c.nA;
    const int D = 4 * nA * nO * nO, nout = 4 * D, ntile = nout / 4;
    const int tt = blockIdx.x * 256 + threadIdx.x;
    if (tt >= ntile) return;
    const long b0 = (long)blockIdx.y * c.chunk;
    const long b1 = b0 + c.chunk < c.np ? b0 + c.chunk : c.np;
    int i, sp, ta, tb;
    tile_of(tt, nO, i, sp, ta, tb);
    const int cg = 1 - (sp & 1), rx = sp >> 1;
    const int a0 = 2 * i * nO + 2 * ta, bq = 2 * i * nO + 2 * tb;
    double2 acc[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = make_double2(0.0, 0.0);
    for (long pl = b0; pl < b1; ++pl) {
        const long p = c.p0 + pl;
        double w;
        if (c.real) {
            const long ie = p / c.nk_re, ik = p - (p / c.nk_re) * c.nk_re;
            w = c.p2[ie] * c.wre[ik];
        } else {
            w = c.wim[p];
        }
        const double2 *Q = raw + (size_t)pl * d * d;
        const double2 *z0 = Zg + (size_t)pl * d * d + (size_t)a0 * d + cg * n, *z1 = z0 + d;
        double2 o00 = make_double2(0.0, 0.0), o01 = o00, o10 = o00, o11 = o00;
        for (int lx = 0; lx < n; ++lx) {
            const int l = lx / nO, x = lx - l * nO, qr = (2 * l + rx) * nO + x;
            const double2 q0 = herm ? zconj(Q[bq * d + qr]) : Q[qr * d + bq];
            const double2 q1 = herm ? zconj(Q[(bq + 1) * d + qr]) : Q[qr * d + bq + 1];
            const double2 za = z0[lx], zb = z1[lx];
            zfma(o00, za, q0);
            zfma(o01, za, q1);
            zfma(o10, zb, q0);
            zfma(o11, zb, q1);
        }
        acc[0].x += w * o00.x; acc[0].y += w * o00.y;
        acc[1].x += w * o01.x; acc[1].y += w * o01.y;
        acc[2].x += w * o10.x; acc[2].y += w * o10.y;
        acc[3].x += w * o11.x; acc[3].y += w * o11.y;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int a = 2 * ta + (e >> 1), b = 2 * tb + (e & 1);
        const int rn = a / nO, nu = a - rn * nO, cm = b / nO, mu = b - cm * nO;
        const int X = 4 * ((i * nO + mu) * nO + nu) + 2 * cm + 1 - rn;
        double2 *o = partial + (size_t)blockIdx.y * nout + X + (size_t)D * sp;
        if (add) {
            o->x += acc[e].x;
            o->y += acc[e].y;
        } else {
            *o = acc[e];
        }
    }
}

// acc[o] += sum over chunks of partial (fixed order)
__global__ void current_sum_kernel(const double2 *partial, long nchunk, int nout, double2 *acc)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    double2 s = acc[o];
    for (long k = 0; k < nchunk; ++k) {
        const double2 v = partial[(size_t)k * nout + o];
        s.x += v.x;
        s.y += v.y;
    }
    acc[o] = s;
}

// out[X + D s'] = sum_Y pre[X + D Y] acc[Y + D s'] / 2 pi (pre = NULL: identity)
__global__ void current_prefactor_kernel(const double2 *pre, const double2 *acc, int D, double2 *out)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= 4 * D) return;
    const int X = o % D, sp = o / D;
    double2 s = make_double2(0.0, 0.0);
    if (pre) {
        for (int Y = 0; Y < D; ++Y) {
            const double2 a = pre[(size_t)X + (size_t)D * Y], b = acc[(size_t)Y + (size_t)D * sp];
            s.x += a.x * b.x - a.y * b.y;
            s.y += a.x * b.y + a.y * b.x;
        }
    } else {
        s = acc[o];
    }
    const double f = 0.15915494309189533577;   // 1 / (2 pi)
    out[o] = make_double2(s.x * f, s.y * f);
}

}  // namespace

extern "C" int th_current_response(th_handle *h, const double *q, double omega, double Ef, double eta, double etap,
                                   int n_x, const double *x2, const double *p2, const double *ue,
                                   const double *prefactor, double *out)
{
    if (!h || !q || !ue || !out || !(eta > 0.0) || !(etap > 0.0) || n_x < 0 || (n_x > 0 && (!x2 || !p2))) {
        th_set_error("th_current_response: bad arguments");
        return TH_E_ARG;
    }
    if (!h->d_kim || h->n_im <= 0) {
        th_set_error("th_set_imag_points must be called before th_current_response");
        return TH_E_STATE;
    }
    const bool real_axis = std::fabs(omega) >= 1e-15 && n_x > 0;
    if (real_axis && (!h->d_kre || h->nk_re <= 0)) {
        th_set_error("th_set_real_kmesh must be called before th_current_response at omega != 0");
        return TH_E_STATE;
    }
    const int d = h->d, D = h->D, nout = 4 * D, n = d / 2;
    int rc = 0;
    if (real_axis && (rc = th_upload_nodes(h, n_x, x2, p2))) return rc;
    // the LDS kernel holds P, Q, Z and T of a pair and D x 4 outputs in
    // registers; otherwise (or by request) the three-stage form
    const size_t lds = sizeof(double2) * ((size_t)3 * d * d + (size_t)n * n + (size_t)std::max(1, h->nhop));
    const bool general = h->current_general || lds > 160 * 1024 || D > TH_CU_RT * TH_CU_NT;
    // pairs per batch: the two raw G (d^2 complex each), and in the three-stage
    // form Z (d^2) and T ((d/2)^2), within the panel budget
    const size_t per = (size_t)(general ? 3 : 2) * d * d * sizeof(double2) + (general ? (size_t)n * n * sizeof(double2) : 0);
    const long n_re = real_axis ? (long)n_x * h->nk_re : 0;
    const long bp = std::max<long>(1, std::min<long>(std::max(h->n_im, n_re), (long)(((size_t)h->batch_mb << 20) / per)));
    const long nch_max = std::min<long>(bp, TH_CU_WG);
    double2 *d_raw = nullptr, *d_part = nullptr, *d_acc = nullptr, *d_pre = nullptr, *d_out = nullptr;
    auto cleanup = [&]() { hipFree(d_raw); hipFree(d_part); hipFree(d_acc); hipFree(d_pre); hipFree(d_out); };
    if (hipMalloc(&d_raw, per * bp) != hipSuccess || hipMalloc(&d_part, sizeof(double2) * nch_max * nout) != hipSuccess ||
        hipMalloc(&d_acc, sizeof(double2) * nout) != hipSuccess || hipMalloc(&d_out, sizeof(double2) * nout) != hipSuccess ||
        (prefactor && hipMalloc(&d_pre, sizeof(double2) * (size_t)D * D) != hipSuccess)) {
        cleanup();
        (void)hipGetLastError();
        th_set_error("th_current_response: out of device memory");
        return TH_E_HIP;
    }
    double2 *d_rawq = d_raw, *d_rawg = d_raw + (size_t)bp * d * d;
    double2 *d_Z = d_rawg + (size_t)bp * d * d, *d_T = d_Z + (size_t)bp * d * d;   // three-stage form only
    const size_t lds_t = sizeof(double2) * (size_t)std::max(1, h->nhop);
    if (general && lds_t > 64 * 1024) {
        cleanup();
        th_set_error("th_current_response: too many hops for the phase table in LDS");
        return TH_E_UNSUPPORTED;
    }
    if (!general && hipFuncSetAttribute((const void *)current_pair_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds) != hipSuccess) {
        cleanup();
        (void)hipGetLastError();
        th_set_error("th_current_response: workgroup memory attribute refused");
        return TH_E_HIP;
    }
    if (!rc) rc = th_ensure_panels(h, bp);
    if (!rc) rc = th_reset_status(h);
    if (!rc) rc = th_launch_zero(h, reinterpret_cast<double *>(d_acc), (size_t)2 * nout);
    if (!rc && prefactor &&
        hipMemcpyAsync(d_pre, prefactor, sizeof(double2) * (size_t)D * D, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        rc = TH_E_HIP;
    current_params cp{};
    cp.nA = h->nA; cp.nO = h->nOrb; cp.d = d; cp.nhop = h->nhop; cp.nk_re = h->nk_re;
    cp.u0 = ue[0]; cp.u1 = ue[1]; cp.u2 = ue[2];
    cp.R = h->d_R; cp.t = h->d_t; cp.hoff = h->d_hoff;
    cp.kim = h->d_kim; cp.wim = h->d_wim; cp.kre = h->d_kre; cp.wre = h->d_wre; cp.p2 = h->d_p2;
    for (int axis = 0; axis < (real_axis ? 2 : 1) && !rc; ++axis) {
        const long total = axis == 0 ? h->n_im : n_re;
        for (long p0 = 0; p0 < total && !rc; p0 += bp) {
            const long np = std::min(bp, total - p0);
            th_green_args a{};
            a.p0 = p0; a.np = np; a.q[0] = q[0]; a.q[1] = q[1]; a.q[2] = q[2];
            a.omega = omega; a.Ef = Ef; a.eta = eta; a.etap = etap; a.prec = TH_PREC_FP64;
            a.mode = axis == 0 ? TH_G_IMAG_A : TH_G_REAL_A;
            a.panel = h->d_pA;
            a.raw = d_rawq;
            if ((rc = th_launch_green(h, a))) break;
            a.mode = axis == 0 ? TH_G_IMAG_B : TH_G_REAL_B;
            a.panel = h->d_pB;
            a.raw = d_rawg;
            if ((rc = th_launch_green(h, a))) break;
            cp.np = np; cp.p0 = p0; cp.real = axis;
            cp.chunk = (np + TH_CU_WG - 1) / TH_CU_WG;
            const long nch = (np + cp.chunk - 1) / cp.chunk;
            if (general) {
                hipLaunchKernelGGL(current_t_kernel, dim3((unsigned)np), dim3(256), lds_t, h->stream, cp, d_T);
                const unsigned gz = (unsigned)(((d / 2) * (d / 2) + 255) / 256), go = (unsigned)((D + 255) / 256);
                for (int term = 0; term < (axis ? 1 : 2); ++term) {
                    hipLaunchKernelGGL(current_z_kernel, dim3((unsigned)np, gz), dim3(256), 0, h->stream, cp,
                                       (const double2 *)d_rawq, (const double2 *)d_T, term, d_Z);
                    hipLaunchKernelGGL(current_o_kernel, dim3(go, (unsigned)nch), dim3(256), 0, h->stream, cp,
                                       (const double2 *)d_Z, (const double2 *)d_rawg, (axis || term) ? 1 : 0,
                                       term, d_part);
                }
            } else {
                hipLaunchKernelGGL(current_pair_kernel, dim3((unsigned)nch), dim3(TH_CU_NT), lds, h->stream, cp,
                                   (const double2 *)d_rawq, (const double2 *)d_rawg, d_part);
            }
            hipLaunchKernelGGL(current_sum_kernel, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, h->stream,
                               (const double2 *)d_part, nch, nout, d_acc);
            if (hipGetLastError() != hipSuccess) { rc = TH_E_HIP; th_set_error("th_current_response: kernel launch"); }
        }
    }
    if (!rc) {
        hipLaunchKernelGGL(current_prefactor_kernel, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, h->stream,
                           (const double2 *)d_pre, (const double2 *)d_acc, D, d_out);
        if (hipGetLastError() != hipSuccess) { rc = TH_E_HIP; th_set_error("th_current_response: kernel launch"); }
    }
    if (!rc && hipMemcpyAsync(out, d_out, sizeof(double2) * nout, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
        rc = TH_E_HIP;
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = TH_E_HIP;
    if (!rc) rc = th_check_status(h);
    cleanup();
    return rc;
}
