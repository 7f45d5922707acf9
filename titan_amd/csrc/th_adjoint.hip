// th_adjoint.hip -- reverse-mode derivative of the chi(omega) path on gfx950.
//
// Forward, per frequency: H(k) from the hops t_n and h0, G = (z - H)^-1 for
// every (k,E) pair, the contraction T = sum_p A_p B_p^T of the G panels,
// chi0 = finalize(T), and chi = (I + chi0 U)^-1 chi0.  The reference has no
// derivatives; this file runs the chain backwards for a real loss L with
// every cotangent written as xhat = dL/dRe x + i dL/dIm x:
//   Dyson      chi0hat = M^-H chihat N^H, Uhat = -chi0^H M^-H chihat chi^H,
//              M = I + chi0 U, N = I - U chi, with M^-H applied through the
//              support S of U (the push-through identity of the forward);
//   finalize   That[b1(X,Y)] = chi0hat[X,Y]/2pi, That[b2(X,Y)] += conj(..)/2pi
//              (imaginary axis), Fhat[b1] = chi0hat/2pi (real axis);
//   contract   Ahat_p = That conj(B_p), Bhat_p = That^T conj(A_p) (MFMA,
//              th_gemm.hip), then Ghat of each Green function from the panel
//              formulas (weights, the real-axis Gq - Gq^H and G^H);
//   inversion  Hhat = G^H Ghat G^H for both Green functions of the pair;
//   assembly   that_n = sum_p exp(-i k_p.R_n) sum_s Hhat_p[(jb,s,nu),(ib,s,mu)],
//              h0hat = sum_p Hhat_p.
// Green functions are recomputed batch by batch (never stored for the
// backward), except the omega-independent imaginary-axis ones G(k, E_F +
// i(y + eta')) that the forward keeps in its panel cache: those are read from
// there.  All sums are FP64 and run in a fixed order, so the result is
// bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "th_internal.h"

namespace {

__device__ inline double2 zmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

inline unsigned nblk(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// Dyson adjoint: dense complex products, column-major.
// C = alpha op(A) op(B) + beta C, op 0 = as is, 1 = conjugate transpose.
// A 16 x 16 tile per 256-thread workgroup, K staged through LDS by 16; the
// loads follow the memory order of each operand.
__global__ __launch_bounds__(256) void zgemm_kernel(int opA, int opB, int m, int n, int k, double2 alpha,
                                                    const double2 *__restrict__ A, int lda,
                                                    const double2 *__restrict__ B, int ldb, double2 beta,
                                                    double2 *__restrict__ C, int ldc)
{
    __shared__ double2 As[16][17], Bs[16][17];
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
    const int r = t & 15, c = t >> 4;
    double2 acc = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < k; k0 += 16) {
        {   // As[row][kk] = op(A)[i0+row, k0+kk]
            int row, kk;
            if (opA == 0) { row = t & 15; kk = t >> 4; } else { kk = t & 15; row = t >> 4; }
            const int gi = i0 + row, gk = k0 + kk;
            double2 v = make_double2(0.0, 0.0);
            if (gi < m && gk < k) v = (opA == 0) ? A[gi + (size_t)lda * gk] : zconj(A[gk + (size_t)lda * gi]);
            As[row][kk] = v;
        }
        {   // Bs[kk][col] = op(B)[k0+kk, j0+col]
            int kk, col;
            if (opB == 0) { kk = t & 15; col = t >> 4; } else { col = t & 15; kk = t >> 4; }
            const int gk = k0 + kk, gj = j0 + col;
            double2 v = make_double2(0.0, 0.0);
            if (gj < n && gk < k) v = (opB == 0) ? B[gk + (size_t)ldb * gj] : zconj(B[gj + (size_t)ldb * gk]);
            Bs[kk][col] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) acc = zadd(acc, zmul(As[r][kk], Bs[kk][c]));
        __syncthreads();
    }
    const int gi = i0 + r, gj = j0 + c;
    if (gi < m && gj < n) {
        double2 *o = C + gi + (size_t)ldc * gj;
        double2 v = zmul(alpha, acc);
        if (beta.x != 0.0 || beta.y != 0.0) v = zadd(v, zmul(beta, *o));
        *o = v;
    }
}

// out[x + D a] = M[x + D S[a]]   (D x r)
__global__ void gather_cols(const double2 *M, int D, const int *S, int r, double2 *out)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * r) return;
    size_t x = e % D, a = e / D;
    out[e] = M[x + (size_t)D * S[a]];
}

// out[a + r y] = M[S[a] + D y]   (r x D)
__global__ void gather_rows(const double2 *M, int D, const int *S, int r, double2 *out)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * r) return;
    size_t a = e % r, y = e / r;
    out[e] = M[S[a] + (size_t)D * y];
}

// X[S[a] + D y] -= Z[a + r y]   (rows of S are distinct: no two threads share an entry)
__global__ void scatter_sub_rows(double2 *X, int D, const int *S, int r, const double2 *Z)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * r) return;
    size_t a = e % r, y = e / r;
    double2 *o = X + S[a] + (size_t)D * y;
    *o = make_double2(o->x - Z[e].x, o->y - Z[e].y);
}

// X[x + D S[b]] -= Q[x + D b]
__global__ void scatter_sub_cols(double2 *X, int D, const int *S, int r, const double2 *Q)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * r) return;
    size_t x = e % D, b = e / D;
    double2 *o = X + x + (size_t)D * S[b];
    *o = make_double2(o->x - Q[e].x, o->y - Q[e].y);
}

// W = I + chi0[S,S] U_S  (r x r, column-major) and its inverse by
// Gauss-Jordan with partial pivoting in one workgroup (r is a few hundred)
__global__ void form_w_adj(const double2 *chi0, int D, const int *S, const double2 *US, int r, double2 *W)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= r * r) return;
    int a = e % r, b = e / r;
    double2 s = make_double2(a == b ? 1.0 : 0.0, 0.0);
    for (int c = 0; c < r; ++c) s = zadd(s, zmul(chi0[S[a] + (size_t)D * S[c]], US[c + (size_t)r * b]));
    W[e] = s;
}

__global__ __launch_bounds__(1024) void gj_inverse(double2 *W, int r, int *perm, int *status)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double2 *colk = reinterpret_cast<double2 *>(smem);
    double2 *prow = colk + r;
    __shared__ double redv[1024];
    __shared__ int redi[1024];
    __shared__ int s_p;
    const int tid = threadIdx.x, NT = blockDim.x;
#define Wm(a, b) W[(a) + (size_t)r * (b)]
    for (int k = 0; k < r; ++k) {
        double v = -1.0;
        int ix = r;
        for (int t = k + tid; t < r; t += NT) {
            double a = fabs(Wm(t, k).x) + fabs(Wm(t, k).y);
            if (a > v) { v = a; ix = t; }
        }
        redv[tid] = v;
        redi[tid] = ix;
        __syncthreads();
        for (int off = NT / 2; off > 0; off >>= 1) {
            if (tid < off) {
                double ov = redv[tid + off];
                int oi = redi[tid + off];
                if (ov > redv[tid] || (ov == redv[tid] && oi < redi[tid])) { redv[tid] = ov; redi[tid] = oi; }
            }
            __syncthreads();
        }
        if (tid == 0) { s_p = redi[0]; perm[k] = redi[0]; }
        __syncthreads();
        const int p = s_p;
        double2 piv = Wm(p, k);
        __syncthreads();
        for (int t = tid; t < r; t += NT) {
            double2 rk = Wm(k, t), rp = Wm(p, t);
            colk[t] = (t == k) ? Wm(p, k) : (t == p ? Wm(k, k) : Wm(t, k));
            prow[t] = rp;
            if (p != k) Wm(p, t) = rk;
        }
        __syncthreads();
        if (piv.x == 0.0 && piv.y == 0.0) {
            if (tid == 0) atomicMin(status, k + 1);
            piv = make_double2(1.0, 0.0);
        }
        const double nn = piv.x * piv.x + piv.y * piv.y;
        const double2 inv = make_double2(piv.x / nn, -piv.y / nn);
        for (int t = tid; t < r; t += NT) {
            double2 pr = (t == k) ? inv : zmul(prow[t], inv);
            prow[t] = pr;
            Wm(k, t) = pr;
        }
        __syncthreads();
        for (size_t e = tid; e < (size_t)r * r; e += NT) {
            int a = (int)(e % r), c = (int)(e / r);
            if (a == k) continue;
            double2 f = colk[a];
            if (c == k) {
                double2 q = zmul(f, inv);
                Wm(a, c) = make_double2(-q.x, -q.y);
            } else {
                double2 q = zmul(f, prow[c]);
                Wm(a, c) = make_double2(Wm(a, c).x - q.x, Wm(a, c).y - q.y);
            }
        }
        __syncthreads();
    }
    for (int k = r - 1; k >= 0; --k) {
        int pk = perm[k];
        if (pk != k)
            for (int a = tid; a < r; a += NT) {
                double2 t0 = Wm(a, k);
                Wm(a, k) = Wm(a, pk);
                Wm(a, pk) = t0;
            }
        __syncthreads();
    }
#undef Wm
}

// ---------------------------------------------------------------------------
// Adjoint of finalize_kernel (th_gemm.hip): chi0 = (T[b1] + conj(T[b2]) +
// F[b1]) / 2pi with the index maps below; b1 and b2 are bijections of the
// chi index pairs onto the unpadded accumulator entries, so each pass
// writes every entry once.
__device__ inline void chi_maps(int X, int Y, int nA, int nOrb, int Mpad, size_t &b1, size_t &b2)
{
    const int bs = 2 * nOrb;
    int s1 = X & 3, r = X >> 2;
    const int nu = r % nOrb; r /= nOrb;
    const int mu = r % nOrb;
    const int i = r / nOrb;
    int s2 = Y & 3;
    r = Y >> 2;
    const int xi = r % nOrb; r /= nOrb;
    const int ga = r % nOrb;
    const int j = r / nOrb;
    const int a1 = (s1 == 0 || s1 == 2) ? 1 : 0, c1 = (s1 >= 2) ? 1 : 0;
    const int a2 = (s2 == 0 || s2 == 2) ? 1 : 0, c2 = (s2 >= 2) ? 1 : 0;
    const size_t MM = (size_t)Mpad * Mpad;
    const int rr = (a1 * nOrb + nu) * bs + (a2 * nOrb + ga);
    const int cc = (c2 * nOrb + xi) * bs + (c1 * nOrb + mu);
    b1 = (size_t)(i * nA + j) * 2 * MM + (size_t)rr * Mpad + cc;
    const int rr2 = (a2 * nOrb + ga) * bs + (a1 * nOrb + nu);
    const int cc2 = (c1 * nOrb + mu) * bs + (c2 * nOrb + xi);
    b2 = (size_t)(j * nA + i) * 2 * MM + (size_t)rr2 * Mpad + cc2;
}

__global__ void unfinalize_1(const double2 *__restrict__ chat, double *__restrict__ acc, int nA, int nOrb, int Mpad,
                             int D, int real_axis)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * D) return;
    size_t b1, b2;
    chi_maps((int)(e % D), (int)(e / D), nA, nOrb, Mpad, b1, b2);
    const double inv2pi = 0.15915494309189533577;
    const size_t MM = (size_t)Mpad * Mpad, axis = (size_t)nA * nA * 2 * MM;
    const double2 v = chat[e];
    acc[b1] = v.x * inv2pi;
    acc[b1 + MM] = v.y * inv2pi;
    if (real_axis) {
        acc[axis + b1] = v.x * inv2pi;
        acc[axis + b1 + MM] = v.y * inv2pi;
    }
}

__global__ void unfinalize_2(const double2 *__restrict__ chat, double *__restrict__ acc, int nA, int nOrb, int Mpad,
                             int D)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * D) return;
    size_t b1, b2;
    chi_maps((int)(e % D), (int)(e / D), nA, nOrb, Mpad, b1, b2);
    const double inv2pi = 0.15915494309189533577;
    const size_t MM = (size_t)Mpad * Mpad;
    const double2 v = chat[e];
    acc[b2] += v.x * inv2pi;
    acc[b2 + MM] -= v.y * inv2pi;
}

// ---------------------------------------------------------------------------
// Ghat (row-major d x d per pair) from the panel cotangent of one side.
struct assemble_params {
    int mode, nA, bs, d, Mpad, nA2;
    long p0, np, nk_re;
    int nx;                  // real-axis nodes; real pairs grouped k-major (th_real_grouped)
    const double *wim, *p2, *wre;
    const double *hat;       // [np][nA2][2][Mpad]
    double2 *out;            // [np][d][d]
};

__global__ void assemble_kernel(assemble_params a)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t dd = (size_t)a.d * a.d;
    if (e >= (size_t)a.np * dd) return;
    const long pl = (long)(e / dd);
    const int rc = (int)(e - pl * dd);
    const int r = rc / a.d, c = rc - (rc / a.d) * a.d;
    const int bs = a.bs, nA = a.nA, Mpad = a.Mpad;
    const int ra = r / bs, ca = c / bs, rr = r - ra * bs, cc = c - ca * bs;
    const double *H = a.hat + (size_t)pl * a.nA2 * 2 * Mpad;
    auto hat = [&](int blk, int idx) { return make_double2(H[(size_t)blk * 2 * Mpad + idx], H[(size_t)blk * 2 * Mpad + Mpad + idx]); };
    const long p = a.p0 + pl;
    double2 v;
    if (a.mode == TH_G_IMAG_A) {            // A = w Gq(i,j):  Gqhat = w Ahat
        const double w = a.wim[p];
        const double2 x = hat(ra * nA + ca, rr * bs + cc);
        v = make_double2(w * x.x, w * x.y);
    } else if (a.mode == TH_G_IMAG_B) {     // B(i,j) = G(j,i)
        v = hat(ca * nA + ra, rr * bs + cc);
    } else if (a.mode == TH_G_REAL_A) {     // A = -i w (Gq - Gq^H):  Gqhat = i w (Ahat + Ahat^H)
        unsigned ie, ik;
        th_real_grouped((unsigned)p, (unsigned)a.nk_re, (unsigned)a.nx, ie, ik);
        const double w = a.p2[ie] * a.wre[ik];
        const double2 x = hat(ra * nA + ca, rr * bs + cc), y = hat(ca * nA + ra, cc * bs + rr);
        const double sr = x.x + y.x, si = x.y - y.y;
        v = make_double2(-w * si, w * sr);
    } else {                                // B(i,j) = G^H(j,i):  Ghat(i,j)[cc][rr] = conj(Bhat(i,j)[rr][cc])
        v = zconj(hat(ra * nA + ca, cc * bs + rr));
    }
    a.out[e] = v;
}

// ---------------------------------------------------------------------------
// Batched C_p = op(A_p) op(B_p), d x d row-major FP64 complex per pair,
// op 1 = conjugate transpose.  A 32 x 32 output tile per 256-thread
// workgroup (2 x 2 entries per thread), K staged through LDS by 16.
__global__ __launch_bounds__(256) void bmm_kernel(const double2 *__restrict__ A, int opA,
                                                  const double2 *__restrict__ B, int opB, double2 *__restrict__ C,
                                                  int d, int ntile)
{
    __shared__ double2 As[32][17], Bs[16][33];
    const int t = threadIdx.x;
    const long pl = blockIdx.x / (ntile * ntile);
    const int tile = (int)(blockIdx.x - pl * ntile * ntile);
    const int i0 = (tile / ntile) * 32, j0 = (tile % ntile) * 32;
    const size_t dd = (size_t)d * d;
    const double2 *Ap = A + pl * dd, *Bp = B + pl * dd;
    const int tx = t & 15, ty = t >> 4;
    double2 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < d; k0 += 16) {
        for (int u = 0; u < 2; ++u) {
            const int v = t + 256 * u;
            int row, kk;
            if (opA == 0) { kk = v & 15; row = v >> 4; } else { row = v & 31; kk = v >> 5; }
            const int gi = i0 + row, gk = k0 + kk;
            double2 x = make_double2(0.0, 0.0);
            if (gi < d && gk < d) x = (opA == 0) ? Ap[(size_t)gi * d + gk] : zconj(Ap[(size_t)gk * d + gi]);
            As[row][kk] = x;
        }
        for (int u = 0; u < 2; ++u) {
            const int v = t + 256 * u;
            int kk, col;
            if (opB == 0) { col = v & 31; kk = v >> 5; } else { kk = v & 15; col = v >> 4; }
            const int gk = k0 + kk, gj = j0 + col;
            double2 x = make_double2(0.0, 0.0);
            if (gj < d && gk < d) x = (opB == 0) ? Bp[(size_t)gk * d + gj] : zconj(Bp[(size_t)gj * d + gk]);
            Bs[kk][col] = x;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const double2 a0 = As[ty][kk], a1 = As[ty + 16][kk];
            const double2 b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
            acc[0][0] = zadd(acc[0][0], zmul(a0, b0));
            acc[0][1] = zadd(acc[0][1], zmul(a0, b1));
            acc[1][0] = zadd(acc[1][0], zmul(a1, b0));
            acc[1][1] = zadd(acc[1][1], zmul(a1, b1));
        }
        __syncthreads();
    }
    double2 *Cp = C + pl * dd;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int gi = i0 + ty + 16 * a, gj = j0 + tx + 16 * b;
            if (gi < d && gj < d) Cp[(size_t)gi * d + gj] = acc[a][b];
        }
}

// ---------------------------------------------------------------------------
// Hhat = G^H Ghat G^H with G and Ghat of a pair in LDS (d a multiple of 3,
// d <= 54).  A thread owns a 3 x 3 block of the d x d result, so each LDS
// value it loads feeds three complex FMAs; ppw pairs share a workgroup of
// ppw (d/3)^2 threads (rounded up to whole waves).  X = Ghat G^H is held in
// registers, then overwrites Ghat in LDS for the second product.  Ghat is
// assembled from the panel cotangent as assemble_kernel does.
__device__ inline double2 ghat_entry(const assemble_params &a, const double *H, int r, int c, double w)
{
    const int bs = a.bs, nA = a.nA, Mpad = a.Mpad;
    const int ra = r / bs, ca = c / bs, rr = r - ra * bs, cc = c - ca * bs;
    auto hat = [&](int blk, int idx) { return make_double2(H[(size_t)blk * 2 * Mpad + idx], H[(size_t)blk * 2 * Mpad + Mpad + idx]); };
    if (a.mode == TH_G_IMAG_A) {
        const double2 x = hat(ra * nA + ca, rr * bs + cc);
        return make_double2(w * x.x, w * x.y);
    }
    if (a.mode == TH_G_IMAG_B) return hat(ca * nA + ra, rr * bs + cc);
    if (a.mode == TH_G_REAL_A) {
        const double2 x = hat(ra * nA + ca, rr * bs + cc), y = hat(ca * nA + ra, cc * bs + rr);
        return make_double2(-w * (x.y - y.y), w * (x.x + y.x));
    }
    return zconj(hat(ra * nA + ca, cc * bs + rr));
}

// Hhat summed over a segment of consecutive pairs on one k-point (such pairs
// share the phases of the hop reduction): one workgroup per segment, ppw pairs
// at a time with their G and Ghat in LDS; a thread owns a 3 x 3 block of one
// pair's result, so each LDS value it loads feeds three complex FMAs, and sums
// it over the pairs it takes; the ppw partial sums are added in a fixed order
// at the end.  X = Ghat G^H is held in registers, then overwrites Ghat in LDS
// for the second product.  seg holds the segments' first pairs (indices of
// the whole list) and the end of the last.  The products run in R (FP32 for
// the mixed path, at twice the FP64 rate and half the LDS traffic), the
// segment sums in FP64.
template <typename R> struct c2of;
template <> struct c2of<double> { typedef double2 T; };
template <> struct c2of<float> { typedef float2 T; };

template <typename R>
__global__ __launch_bounds__(512) void hhat_seg_kernel(assemble_params a, const double2 *__restrict__ Graw,
                                                       const int *__restrict__ seg, int ppw, double2 *__restrict__ out)
{
    typedef typename c2of<R>::T c2;
    extern __shared__ __align__(16) unsigned char smem[];
    const int d = a.d, dd = d * d, nb3 = d / 3, nblk3 = nb3 * nb3;
    const int tid = threadIdx.x, NT = blockDim.x;
    const long q0 = seg[blockIdx.x] - a.p0, q1 = seg[blockIdx.x + 1] - a.p0;   // the batch's pairs
    const int s = tid / nblk3, bidx = tid - s * nblk3;
    const int r0 = 3 * (bidx / nb3), c0 = 3 * (bidx - (bidx / nb3) * nb3);
    c2 *G = reinterpret_cast<c2 *>(smem) + (size_t)(s < ppw ? s : 0) * 2 * dd;
    c2 *Gh = G + dd;
    double2 tot[3][3];
    c2 acc[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tot[i][j] = make_double2(0.0, 0.0);
    for (long g0 = q0; g0 < q1; g0 += ppw) {
        __syncthreads();                         // the previous group's reads of G, Gh done
        for (int e = tid; e < ppw * dd; e += NT) {
            const int sl = e / dd, f = e - sl * dd;
            const long pl = g0 + sl;
            c2 *Gs = reinterpret_cast<c2 *>(smem) + (size_t)sl * 2 * dd;
            if (pl >= q1) {
                Gs[f] = c2{R(0), R(0)};
                Gs[dd + f] = c2{R(0), R(0)};
                continue;
            }
            const long p = a.p0 + pl;
            double w = 0.0;
            if (a.mode == TH_G_IMAG_A) w = a.wim[p];
            if (a.mode == TH_G_REAL_A) {
                unsigned ie, ik;
                th_real_grouped((unsigned)p, (unsigned)a.nk_re, (unsigned)a.nx, ie, ik);
                w = a.p2[ie] * a.wre[ik];
            }
            const double *H = a.hat + (size_t)pl * a.nA2 * 2 * a.Mpad;
            const double2 gv = Graw[(size_t)pl * dd + f];
            const double2 hv = ghat_entry(a, H, f / d, f - (f / d) * d, w);
            Gs[f] = c2{(R)gv.x, (R)gv.y};
            Gs[dd + f] = c2{(R)hv.x, (R)hv.y};
        }
        __syncthreads();
        const bool active = s < ppw && g0 + s < q1;
        // X = Ghat G^H: X[r][c] = sum_m Ghat[r][m] conj(G[c][m])
        if (active) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) acc[i][j] = c2{R(0), R(0)};
            for (int m = 0; m < d; ++m) {
                c2 x[3], y[3];
                for (int i = 0; i < 3; ++i) x[i] = Gh[(r0 + i) * d + m];
                for (int j = 0; j < 3; ++j) y[j] = G[(c0 + j) * d + m];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        acc[i][j] = c2{fma(x[i].x, y[j].x, fma(x[i].y, y[j].y, acc[i][j].x)),
                                       fma(x[i].y, y[j].x, fma(-x[i].x, y[j].y, acc[i][j].y))};
            }
        }
        __syncthreads();
        if (active)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Gh[(r0 + i) * d + c0 + j] = acc[i][j];
        __syncthreads();
        // Hhat = G^H X: Hhat[r][c] = sum_m conj(G[m][r]) X[m][c], added to the sum
        if (active) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) acc[i][j] = c2{R(0), R(0)};
            for (int m = 0; m < d; ++m) {
                c2 x[3], y[3];
                for (int i = 0; i < 3; ++i) x[i] = G[m * d + r0 + i];
                for (int j = 0; j < 3; ++j) y[j] = Gh[m * d + c0 + j];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        acc[i][j] = c2{fma(x[i].x, y[j].x, fma(x[i].y, y[j].y, acc[i][j].x)),
                                       fma(x[i].x, y[j].y, fma(-x[i].y, y[j].x, acc[i][j].y))};
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    tot[i][j] = make_double2(tot[i][j].x + (double)acc[i][j].x, tot[i][j].y + (double)acc[i][j].y);
        }
    }
    __syncthreads();
    // the ppw partial sums (in the staging space: ppw dd FP64 values fit in the
    // ppw 2 dd staged ones of either type), added in order of s
    double2 *red = reinterpret_cast<double2 *>(smem);
    if (s < ppw)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) red[(size_t)s * dd + (r0 + i) * d + c0 + j] = tot[i][j];
    __syncthreads();
    for (int f = tid; f < dd; f += NT) {
        double2 v = red[f];
        for (int t = 1; t < ppw; ++t) v = zadd(v, red[(size_t)t * dd + f]);
        out[(size_t)blockIdx.x * dd + f] = v;
    }
}

// The same with the order D known at compile time (D = 18, 36, 54: one to
// three atoms of nine spin orbitals).  A thread stages the nine entries
// f = bidx + i (D/3)^2 of its own pair slot, whose offsets in the cotangent
// panel are fixed for the launch and computed once, with its pair's weight
// once per group and all nine entries' global loads issued together; the
// product loops unroll with immediate LDS offsets.
template <typename R, int D>
__global__ __launch_bounds__(512) void hhat_seg_d_kernel(assemble_params a, const double2 *__restrict__ Graw,
                                                         const int *__restrict__ seg, int ppw,
                                                         double2 *__restrict__ out)
{
    typedef typename c2of<R>::T c2;
    constexpr int DD = D * D, NB3 = D / 3, NBLK = NB3 * NB3, NE = DD / NBLK;
    static_assert(NE == 9, "nine staged entries per thread");
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const long q0 = seg[blockIdx.x] - a.p0, q1 = seg[blockIdx.x + 1] - a.p0;   // the batch's pairs
    const int s = tid / NBLK, bidx = tid - s * NBLK;
    const bool slot = s < ppw;
    const int r0 = 3 * (bidx / NB3), c0 = 3 * (bidx - (bidx / NB3) * NB3);
    c2 *G = reinterpret_cast<c2 *>(smem) + (size_t)(slot ? s : 0) * 2 * DD;
    c2 *Gh = G + DD;
    const int mode = a.mode, bs = a.bs, nA = a.nA, Mp = a.Mpad, M2 = 2 * a.Mpad;
    const float rbs = 1.0f / (float)bs;   // (x + 0.5) / bs truncates to x / bs for x < D <= 54
    double2 tot[3][3];
    c2 acc[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) tot[i][j] = make_double2(0.0, 0.0);
    for (long g0 = q0; g0 < q1; g0 += ppw) {
        const long pl = g0 + s;
        const bool active = slot && pl < q1;
        __syncthreads();                         // the previous group's reads of G, Gh done
        if (active) {
            const int p = (int)(a.p0 + pl);
            double w = 1.0;
            if (mode == TH_G_IMAG_A) w = a.wim[p];
            if (mode == TH_G_REAL_A) {
                unsigned ie, ik;
                th_real_grouped((unsigned)p, (unsigned)a.nk_re, (unsigned)a.nx, ie, ik);
                w = a.p2[ie] * a.wre[ik];
            }
            const double *H = a.hat + (size_t)pl * a.nA2 * M2;
            const double2 *Gp = Graw + (size_t)pl * DD;
            // three chunks of three entries: the loads of a chunk in flight together
#pragma unroll 1
            for (int i0 = 0; i0 < NE; i0 += 3) {
                double2 gv[3], hv[3], yv[3];
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    const int f = bidx + (i0 + u) * NBLK, r = f / D, c = f - (f / D) * D;
                    const int ra = (int)(((float)r + 0.5f) * rbs), ca = (int)(((float)c + 0.5f) * rbs);
                    const int rr = r - ra * bs, cc = c - ca * bs;
                    // block (i,j) entry [rr][cc], its transpose (j,i) [cc][rr]
                    const int ox = mode == TH_G_IMAG_B ? (ca * nA + ra) * M2 + rr * bs + cc
                                 : mode == TH_G_REAL_B ? (ra * nA + ca) * M2 + cc * bs + rr
                                 : (ra * nA + ca) * M2 + rr * bs + cc;
                    const int oy = (ca * nA + ra) * M2 + cc * bs + rr;
                    gv[u] = Gp[f];
                    hv[u] = make_double2(H[ox], H[ox + Mp]);
                    yv[u] = mode == TH_G_REAL_A ? make_double2(H[oy], H[oy + Mp]) : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    double2 v = hv[u];
                    if (mode == TH_G_REAL_A)             // Gqhat = i w (Ahat + Ahat^H)
                        v = make_double2(-w * (v.y - yv[u].y), w * (v.x + yv[u].x));
                    else if (mode == TH_G_IMAG_A)        // Gqhat = w Ahat
                        v = make_double2(w * v.x, w * v.y);
                    else if (mode == TH_G_REAL_B)        // Ghat(i,j)[cc][rr] = conj(Bhat(i,j)[rr][cc])
                        v.y = -v.y;
                    const int f = bidx + (i0 + u) * NBLK;
                    G[f] = c2{(R)gv[u].x, (R)gv[u].y};
                    Gh[f] = c2{(R)v.x, (R)v.y};
                }
            }
        }
        __syncthreads();
        // X = Ghat G^H: X[r][c] = sum_m Ghat[r][m] conj(G[c][m])
        if (active) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) acc[i][j] = c2{R(0), R(0)};
#pragma unroll 6
            for (int m = 0; m < D; ++m) {
                c2 x[3], y[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) x[i] = Gh[(r0 + i) * D + m];
#pragma unroll
                for (int j = 0; j < 3; ++j) y[j] = G[(c0 + j) * D + m];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        acc[i][j] = c2{fma(x[i].x, y[j].x, fma(x[i].y, y[j].y, acc[i][j].x)),
                                       fma(x[i].y, y[j].x, fma(-x[i].x, y[j].y, acc[i][j].y))};
            }
        }
        __syncthreads();
        if (active)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) Gh[(r0 + i) * D + c0 + j] = acc[i][j];
        __syncthreads();
        // Hhat = G^H X: Hhat[r][c] = sum_m conj(G[m][r]) X[m][c], added to the sum
        if (active) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) acc[i][j] = c2{R(0), R(0)};
#pragma unroll 6
            for (int m = 0; m < D; ++m) {
                c2 x[3], y[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) x[i] = G[m * D + r0 + i];
#pragma unroll
                for (int j = 0; j < 3; ++j) y[j] = Gh[m * D + c0 + j];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        acc[i][j] = c2{fma(x[i].x, y[j].x, fma(x[i].y, y[j].y, acc[i][j].x)),
                                       fma(x[i].x, y[j].y, fma(-x[i].y, y[j].x, acc[i][j].y))};
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    tot[i][j] = make_double2(tot[i][j].x + (double)acc[i][j].x, tot[i][j].y + (double)acc[i][j].y);
        }
    }
    __syncthreads();
    // the ppw partial sums, added in order of s (threads of slots never
    // active hold zeros)
    double2 *red = reinterpret_cast<double2 *>(smem);
    if (slot)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) red[(size_t)s * DD + (r0 + i) * D + c0 + j] = tot[i][j];
    __syncthreads();
    for (int f = tid; f < DD; f += blockDim.x) {
        double2 v = red[f];
        for (int t = 1; t < ppw; ++t) v = zadd(v, red[(size_t)t * DD + f]);
        out[(size_t)blockIdx.x * DD + f] = v;
    }
}

// Spectral form of the segment sums.  With G_p = V diag(g_p) V^H, g_p,m =
// 1/(z_p - lambda_m), and V the eigenvectors of the segment's k-point,
//   sum_p G_p^H Ghat_p G_p^H = V [ sum_p (conj g_p conj g_p^T) o (V^H Ghat_p V) ] V^H,
// so no pair's G is formed or read: the workgroup stages V and lambda once,
// per pair only Ghat_p and conj g_p.  Y_p = V^H (Ghat_p V) runs in R with a
// thread's 3 x 3 register block as in hhat_seg_d_kernel; W sums the scaled
// Y_p in FP64, the slots' W are added in order of slot, and the result is
// added to the k-point's eigenbasis sum Wk[k] (segments of one k-point share V
// and the phases, whichever axis they belong to); hhat_final_kernel forms
// V Wk V^H once per k-point after the last batch.  One workgroup per segment of
// the list sidx, whose segments have distinct k-points.
struct hspec_params {
    const double2 *V;        // [nk][D][D] row-major, V[r][m] = component r of eigenvector m
    const double *L;         // [nk][D]
    const int *kid;          // unique k-point of an imaginary pair, or of a real-axis mesh point
    const double *yim, *x2;
    double er0, ei0;         // z = er0 + x2[ie] (real axis) + i (ei0 + yim[p] (imaginary axis))
};

template <typename R, int D>
__global__ __launch_bounds__(512) void hhat_spec_kernel(assemble_params a, hspec_params sp,
                                                        const int *__restrict__ seg, const int *__restrict__ sidx,
                                                        int ppw, double2 *__restrict__ Wk)
{
    typedef typename c2of<R>::T c2;
    constexpr int DD = D * D, NB3 = D / 3, NBLK = NB3 * NB3, NE = DD / NBLK;
    static_assert(NE == 9, "nine staged entries per thread");
    // LDS: V (R) | W (FP64) | lambda | conj g per slot (FP64) | Ghat, then X, per slot (R)
    extern __shared__ __align__(16) unsigned char smem[];
    c2 *Vs = reinterpret_cast<c2 *>(smem);
    double2 *Ws = reinterpret_cast<double2 *>(Vs + DD);
    double *lam = reinterpret_cast<double *>(Ws + DD);
    double2 *gs = reinterpret_cast<double2 *>(lam + D);
    c2 *slots = reinterpret_cast<c2 *>(gs + (size_t)ppw * D);
    const int tid = threadIdx.x;
    const bool imag = a.mode == TH_G_IMAG_A || a.mode == TH_G_IMAG_B;
    const int sg = sidx[blockIdx.x];                                   // the segment in the batch
    const int pfirst = seg[sg];
    const long q0 = pfirst - a.p0, q1 = seg[sg + 1] - a.p0;             // the batch's pairs
    const int s = tid / NBLK, bidx = tid - s * NBLK;
    const bool slot = s < ppw;
    const int r0 = 3 * (bidx / NB3), c0 = 3 * (bidx - (bidx / NB3) * NB3);
    c2 *Gh = slots + (size_t)(slot ? s : 0) * DD;
    double2 *gp = gs + (size_t)(slot ? s : 0) * D;
    const int mode = a.mode, bs = a.bs, nA = a.nA, Mp = a.Mpad, M2 = 2 * a.Mpad;
    const float rbs = 1.0f / (float)bs;   // (x + 0.5) / bs truncates to x / bs for x < D <= 54
    // the segment's k-point (all its pairs share it)
    int kb;
    if (imag) {
        kb = sp.kid[pfirst];
    } else {
        unsigned ie, ik;
        th_real_grouped((unsigned)pfirst, (unsigned)a.nk_re, (unsigned)a.nx, ie, ik);
        kb = sp.kid[ik];
    }
    const double2 *Vk = sp.V + (size_t)kb * DD;
    for (int e = tid; e < DD; e += blockDim.x) {
        const double2 v = Vk[e];
        Vs[e] = c2{(R)v.x, (R)v.y};
    }
    for (int m = tid; m < D; m += blockDim.x) lam[m] = sp.L[(size_t)kb * D + m];
    double2 tot[3][3];
    c2 acc[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) tot[i][j] = make_double2(0.0, 0.0);
    for (long g0 = q0; g0 < q1; g0 += ppw) {
        const long pl = g0 + s;
        const bool active = slot && pl < q1;
        __syncthreads();                         // V, lambda staged; the previous group's reads done
        if (active) {
            const int p = (int)(a.p0 + pl);
            double w = 1.0, er = sp.er0, ei = sp.ei0;
            if (imag) {
                ei += sp.yim[p];
                if (mode == TH_G_IMAG_A) w = a.wim[p];
            } else {
                unsigned ie, ik;
                th_real_grouped((unsigned)p, (unsigned)a.nk_re, (unsigned)a.nx, ie, ik);
                er += sp.x2[ie];
                if (mode == TH_G_REAL_A) w = a.p2[ie] * a.wre[ik];
            }
            if (bidx < D) {                      // conj g_m = conj(1 / (z - lambda_m))
                const double x = er - lam[bidx], n = 1.0 / (x * x + ei * ei);
                gp[bidx] = make_double2(x * n, ei * n);
            }
            const double *H = a.hat + (size_t)pl * a.nA2 * M2;
            // offsets in the pair's cotangent: ox the entry read, oy the
            // transposed one of the real-axis A side; one atom of nine spin
            // orbitals (D = 18, Mpad = 336): block (0,0), entry [r][c] at r D + c
            auto stage = [&](auto one_atom) {
                constexpr bool ONE = decltype(one_atom)::value;
                const int mp = ONE ? 336 : Mp;
#pragma unroll 1
                for (int i0 = 0; i0 < NE; i0 += 3) {
                    double2 hv[3], yv[3];
#pragma unroll
                    for (int u = 0; u < 3; ++u) {
                        const int f = bidx + (i0 + u) * NBLK, r = f / D, c = f - (f / D) * D;
                        int ox, oy;
                        if constexpr (ONE) {
                            oy = c * D + r;
                            ox = mode == TH_G_REAL_B ? oy : f;
                        } else {
                            const int ra = (int)(((float)r + 0.5f) * rbs), ca = (int)(((float)c + 0.5f) * rbs);
                            const int rr = r - ra * bs, cc = c - ca * bs;
                            ox = mode == TH_G_IMAG_B ? (ca * nA + ra) * M2 + rr * bs + cc
                               : mode == TH_G_REAL_B ? (ra * nA + ca) * M2 + cc * bs + rr
                               : (ra * nA + ca) * M2 + rr * bs + cc;
                            oy = (ca * nA + ra) * M2 + cc * bs + rr;
                        }
                        hv[u] = make_double2(H[ox], H[ox + mp]);
                        yv[u] = mode == TH_G_REAL_A ? make_double2(H[oy], H[oy + mp]) : make_double2(0.0, 0.0);
                    }
#pragma unroll
                    for (int u = 0; u < 3; ++u) {
                        double2 v = hv[u];
                        if (mode == TH_G_REAL_A)             // Gqhat = i w (Ahat + Ahat^H)
                            v = make_double2(-w * (v.y - yv[u].y), w * (v.x + yv[u].x));
                        else if (mode == TH_G_IMAG_A)        // Gqhat = w Ahat
                            v = make_double2(w * v.x, w * v.y);
                        else if (mode == TH_G_REAL_B)        // Ghat(i,j)[cc][rr] = conj(Bhat(i,j)[rr][cc])
                            v.y = -v.y;
                        Gh[bidx + (i0 + u) * NBLK] = c2{(R)v.x, (R)v.y};
                    }
                }
            };
            if (D == 18 && nA == 1 && Mp == 336)
                stage(std::true_type{});
            else
                stage(std::false_type{});
        }
        __syncthreads();
        // X = Ghat V: X[r][c] = sum_m Ghat[r][m] V[m][c]
        if (active) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) acc[i][j] = c2{R(0), R(0)};
#pragma unroll 6
            for (int m = 0; m < D; ++m) {
                c2 x[3], y[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) x[i] = Gh[(r0 + i) * D + m];
#pragma unroll
                for (int j = 0; j < 3; ++j) y[j] = Vs[m * D + c0 + j];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        acc[i][j] = c2{fma(x[i].x, y[j].x, fma(-x[i].y, y[j].y, acc[i][j].x)),
                                       fma(x[i].x, y[j].y, fma(x[i].y, y[j].x, acc[i][j].y))};
            }
        }
        __syncthreads();
        if (active)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) Gh[(r0 + i) * D + c0 + j] = acc[i][j];
        __syncthreads();
        // Y = V^H X: Y[r][c] = sum_m conj(V[m][r]) X[m][c]; W += conj(g_r) conj(g_c) Y
        if (active) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) acc[i][j] = c2{R(0), R(0)};
#pragma unroll 6
            for (int m = 0; m < D; ++m) {
                c2 x[3], y[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) x[i] = Vs[m * D + r0 + i];
#pragma unroll
                for (int j = 0; j < 3; ++j) y[j] = Gh[m * D + c0 + j];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        acc[i][j] = c2{fma(x[i].x, y[j].x, fma(x[i].y, y[j].y, acc[i][j].x)),
                                       fma(x[i].x, y[j].y, fma(-x[i].y, y[j].x, acc[i][j].y))};
            }
            double2 gr[3], gc[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) { gr[i] = gp[r0 + i]; gc[i] = gp[c0 + i]; }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double fr = fma(gr[i].x, gc[j].x, -gr[i].y * gc[j].y);
                    const double fi = fma(gr[i].x, gc[j].y, gr[i].y * gc[j].x);
                    const double yr = (double)acc[i][j].x, yi = (double)acc[i][j].y;
                    tot[i][j] = make_double2(fma(fr, yr, fma(-fi, yi, tot[i][j].x)), fma(fr, yi, fma(fi, yr, tot[i][j].y)));
                }
        }
    }
    // W = the slots' sums, added in order of slot (slots never active hold zeros)
    for (int t = 0; t < ppw; ++t) {
        __syncthreads();
        if (s == t)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double2 &o = Ws[(r0 + i) * D + c0 + j];
                    o = t == 0 ? tot[i][j] : zadd(o, tot[i][j]);
                }
    }
    __syncthreads();
    // added to the k-point's sum (no other workgroup of the launch has this
    // k-point; launches follow the segments' order)
    double2 *o = Wk + (size_t)kb * DD;
    for (int e = tid; e < DD; e += blockDim.x) o[e] = zadd(o[e], Ws[e]);
}

// Hhat_k = V_k W_k V_k^H for k-points k0 .. k0 + gridDim.x - 1 (the sums the
// spectral kernel accumulated in the eigenbasis), FP64, into out[k - k0]
template <int D>
__global__ __launch_bounds__(256) void hhat_final_kernel(const double2 *__restrict__ Wk, const double2 *__restrict__ V,
                                                         long k0, double2 *__restrict__ out)
{
    constexpr int DD = D * D, NU = (DD + 255) / 256;
    extern __shared__ __align__(16) unsigned char smem[];   // W | V, 2 D^2 FP64 complex
    double2 *Ws = reinterpret_cast<double2 *>(smem), *Vs = Ws + DD;
    const long k = k0 + blockIdx.x;
    const int tid = threadIdx.x;
    for (int e = tid; e < DD; e += 256) {
        Ws[e] = Wk[(size_t)k * DD + e];
        Vs[e] = V[(size_t)k * DD + e];
    }
    __syncthreads();
    double2 u[NU];
#pragma unroll
    for (int q = 0; q < NU; ++q) {
        const int e = tid + 256 * q, r = e / D, c = e - (e / D) * D;
        double2 v = make_double2(0.0, 0.0);
        if (e < DD)
#pragma unroll 3
            for (int m = 0; m < D; ++m) {   // v += W[r][m] conj(V[c][m])
                const double2 x = Ws[r * D + m], y = Vs[c * D + m];
                v = make_double2(fma(x.x, y.x, fma(x.y, y.y, v.x)), fma(x.y, y.x, fma(-x.x, y.y, v.y)));
            }
        u[q] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NU; ++q) {
        const int e = tid + 256 * q;
        if (e < DD) Ws[e] = u[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NU; ++q) {
        const int e = tid + 256 * q, r = e / D, c = e - (e / D) * D;
        if (e < DD) {
            double2 v = make_double2(0.0, 0.0);
#pragma unroll 3
            for (int m = 0; m < D; ++m) {   // v += V[r][m] U[m][c]
                const double2 x = Vs[r * D + m], y = Ws[m * D + c];
                v = make_double2(fma(x.x, y.x, fma(-x.y, y.y, v.x)), fma(x.x, y.y, fma(x.y, y.x, v.y)));
            }
            out[(size_t)blockIdx.x * DD + e] = v;
        }
    }
}

// LDS bytes of hhat_spec_kernel
inline size_t hspec_lds(int D, int ppw, bool f32)
{
    const size_t c2 = f32 ? sizeof(float2) : sizeof(double2), dd = (size_t)D * D;
    return dd * (c2 + sizeof(double2)) + sizeof(double) * D + sizeof(double2) * ppw * D + c2 * ppw * dd;
}

// Hhat of the batch's pairs summed over each segment (the generic path)
__global__ void seg_sum_kernel(const double2 *__restrict__ Hh, const int *__restrict__ seg, long p0, int nseg, int dd,
                               double2 *__restrict__ out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nseg * dd) return;
    const int sg = (int)(e / dd), f = (int)(e - (size_t)sg * dd);
    double2 v = make_double2(0.0, 0.0);
    for (long pl = seg[sg] - p0; pl < seg[sg + 1] - p0; ++pl) v = zadd(v, Hh[(size_t)pl * dd + f]);
    out[e] = v;
}

// raw G of the imaginary-axis B side from the omega-independent panel cache:
// B(i,j) = vec G(j,i), so G[(j bs + rr) d + i bs + cc] = B[p][i nA + j][rr bs + cc]
template <typename P>
__global__ void raw_from_bpanel(const P *__restrict__ panel, long np, int nA, int bs, int Mpad, th_rowmap rows,
                                float unsplit, double2 *raw)
{
    const int d = nA * bs;
    const size_t dd = (size_t)d * d;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)np * dd) return;
    const long pl = (long)(e / dd);
    const int f = (int)(e - pl * dd);
    const int r = f / d, c = f - (f / d) * d;
    const int j = r / bs, rr = r - j * bs, i = c / bs, cc = c - i * bs;
    const P *b = panel + ((size_t)th_map_row(rows.pg0 + pl, rows) * nA * nA + (size_t)i * nA + j) * 2 * Mpad;
    if constexpr (sizeof(P) == sizeof(float)) {
        if (unsplit != 0.0f) {   // split words: (hi + lo) 2^-e
            auto dec = [unsplit](P v) {
                const unsigned wv = __builtin_bit_cast(unsigned, (float)v);
                return (double)(((float)__builtin_bit_cast(_Float16, (unsigned short)(wv & 0xffffu)) +
                                 (float)__builtin_bit_cast(_Float16, (unsigned short)(wv >> 16))) * unsplit);
            };
            raw[e] = make_double2(dec(b[rr * bs + cc]), dec(b[Mpad + rr * bs + cc]));
            return;
        }
    }
    raw[e] = make_double2((double)b[rr * bs + cc], (double)b[Mpad + rr * bs + cc]);
}

// exp(-i k.R_n) for the segments of a batch (k of the segment's first pair,
// seg[sg] in the whole list; real pairs grouped k-major over nk k-points and
// nx nodes): ph[sg * nhop + n]
__global__ void phase_kernel(const int *seg, int nseg, int nhop, int imag, int side_q, int nk, int nx, double q0,
                             double q1, double q2, const double *kim, const double *kre, const double *R, double2 *ph)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nseg * nhop) return;
    const int sg = (int)(e / nhop);
    const int n = (int)(e - (size_t)sg * nhop);
    const long p = seg[sg];
    unsigned ie, ik;
    th_real_grouped((unsigned)p, (unsigned)nk, (unsigned)nx, ie, ik);
    const double *k = imag ? kim + 3 * p : kre + 3 * (size_t)ik;
    double k0 = k[0], k1 = k[1], k2 = k[2];
    if (side_q) { k0 += q0; k1 += q1; k2 += q2; }
    double sn, cs;
    sincos(k0 * R[3 * n] + k1 * R[3 * n + 1] + k2 * R[3 * n + 2], &sn, &cs);
    ph[e] = make_double2(cs, -sn);
}

// exp(-i k.R_n) for the unique k-points k0 .. k0 + nk - 1 (plus q on the q side)
__global__ void phase_k_kernel(const double *kuniq, long k0, int nk, int nhop, int side_q, double q0, double q1,
                               double q2, const double *R, double2 *ph)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nk * nhop) return;
    const int i = (int)(e / nhop), n = (int)(e - (size_t)i * nhop);
    const double *k = kuniq + 3 * (size_t)(k0 + i);
    double k0_ = k[0], k1 = k[1], k2 = k[2];
    if (side_q) { k0_ += q0; k1 += q1; k2 += q2; }
    double sn, cs;
    sincos(k0_ * R[3 * n] + k1 * R[3 * n + 1] + k2 * R[3 * n + 2], &sn, &cs);
    ph[e] = make_double2(cs, -sn);
}

// ---------------------------------------------------------------------------
// Partial sums over a chunk of pairs of the hop and h0 cotangents:
//   that[n][nu + nOrb mu] = sum_p exp(-i k_p.R_n) sum_s Hhat_p[(2jb+s)nO+nu][(2ib+s)nO+mu]
//   h0hat[r d + c]        = sum_p Hhat_p[r][c]   (row-major; the host transposes)
// Output e < nhop nO^2 is a hop entry, the rest the h0 entries.  The hop
// entries read the spin-summed blocks S_p, so consecutive outputs read
// consecutive memory.
struct hreduce_params {
    int side_q;              // 1: k + q (the Gq side)
    int imag;                // 1: imaginary-axis pair list
    long p0, np, nk_re;
    int chunk, nout, nhop, nOrb, d;
    double q0, q1, q2;
    const double *kim, *kre, *R;
    const int *jb, *ib;
    int nA;
    const double2 *Hh;       // [np][d][d]
    const double2 *S;        // [np][nA nA][nOrb nOrb] spin-summed blocks of Hh
    const double2 *ph;       // [np][nhop] exp(-i k_p.R_n)
    double2 *partial;        // [nchunk][nout]
};

// Hop entries: a thread owns one entry loc of the spin-summed block of a
// group of at most HG hops of one atom pair (hops are stored grouped by atom
// pair), reads S once per pair and updates HG accumulators.
constexpr int HG = 8;

// A workgroup = 64 outputs x 4 waves; wave y takes the pairs pl = c0 + y,
// c0 + y + 4, ... of the chunk and the four wave sums are added in a fixed
// order through LDS.
__global__ __launch_bounds__(256) void hop_reduce_kernel(hreduce_params a, const int *__restrict__ groups,
                                                         int ngroups)
{
    __shared__ double2 red[4][64][HG];
    const int nO2 = a.nOrb * a.nOrb;
    const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
    const int t = blockIdx.x * 64 + x;
    const bool on = t < ngroups * nO2;
    const int g = on ? t / nO2 : 0, loc = on ? t - g * nO2 : 0;
    const int b = groups[3 * g], h0 = groups[3 * g + 1], cnt = on ? groups[3 * g + 2] : 0;
    const long c0 = (long)blockIdx.y * a.chunk, c1 = min(a.np, c0 + a.chunk);
    const size_t sstride = (size_t)a.nA * a.nA * nO2;
    double2 acc[HG];
#pragma unroll
    for (int c = 0; c < HG; ++c) acc[c] = make_double2(0.0, 0.0);
    if (on)
        for (long pl = c0 + y; pl < c1; pl += 4) {
            const double2 sv = a.S[pl * sstride + (size_t)b * nO2 + loc];
            const double2 *ph = a.ph + pl * a.nhop + h0;
#pragma unroll
            for (int c = 0; c < HG; ++c)
                if (c < cnt) acc[c] = zadd(acc[c], zmul(sv, ph[c]));
        }
#pragma unroll
    for (int c = 0; c < HG; ++c) red[y][x][c] = acc[c];
    __syncthreads();
    if (y != 0 || !on) return;
    double2 *out = a.partial + (size_t)blockIdx.y * a.nout;
#pragma unroll
    for (int c = 0; c < HG; ++c)
        if (c < cnt) {
            double2 v = red[0][x][c];
            v = zadd(v, red[1][x][c]);
            v = zadd(v, red[2][x][c]);
            v = zadd(v, red[3][x][c]);
            out[(size_t)(h0 + c) * nO2 + loc] = v;
        }
}

// h0 entries (row-major here; the host transposes), same work split
__global__ __launch_bounds__(256) void h0_reduce_kernel(hreduce_params a)
{
    __shared__ double2 red[4][64];
    const int dd = a.d * a.d;
    const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
    const int f = blockIdx.x * 64 + x;
    const long c0 = (long)blockIdx.y * a.chunk, c1 = min(a.np, c0 + a.chunk);
    double2 s = make_double2(0.0, 0.0);
    if (f < dd)
        for (long pl = c0 + y; pl < c1; pl += 4) s = zadd(s, a.Hh[pl * (size_t)dd + f]);
    red[y][x] = s;
    __syncthreads();
    if (y != 0 || f >= dd) return;
    double2 v = red[0][x];
    v = zadd(v, red[1][x]);
    v = zadd(v, red[2][x]);
    v = zadd(v, red[3][x]);
    a.partial[(size_t)blockIdx.y * a.nout + (size_t)a.nhop * a.nOrb * a.nOrb + f] = v;
}

// S[pl][j nA + i][nu + nOrb mu] = sum over spin s of Hhat[(2j+s)nOrb+nu][(2i+s)nOrb+mu]
__global__ void spin_sum_kernel(long np, int nA, int nOrb, const double2 *__restrict__ Hh, double2 *__restrict__ S)
{
    const int nO2 = nOrb * nOrb, per = nA * nA * nO2, d = 2 * nOrb * nA;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)np * per) return;
    const long pl = (long)(e / per);
    const int f = (int)(e - (size_t)pl * per);
    const int b = f / nO2, loc = f - b * nO2;
    const int j = b / nA, i = b - j * nA, nu = loc % nOrb, mu = loc / nOrb;
    const double2 *Hp = Hh + (size_t)pl * d * d;
    S[e] = zadd(Hp[(size_t)(2 * j * nOrb + nu) * d + 2 * i * nOrb + mu],
                Hp[(size_t)((2 * j + 1) * nOrb + nu) * d + (2 * i + 1) * nOrb + mu]);
}

// acc[e] += sum_ch partial[ch][e], fixed order
// acc[e] += sum over the nch chunks of partial[ch][e]: one wave per output,
// lane l takes the chunks l, l + 64, ... and the lanes' sums are combined by a
// fixed butterfly (the same order on every call)
__global__ __launch_bounds__(256) void hsum_kernel(const double2 *partial, int nch, int nout, double2 *acc)
{
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (e >= nout) return;
    double2 s = make_double2(0.0, 0.0);
    for (int ch = lane; ch < nch; ch += 64) s = zadd(s, partial[(size_t)ch * nout + e]);
    for (int o = 32; o > 0; o >>= 1) s = make_double2(s.x + __shfl_xor(s.x, o), s.y + __shfl_xor(s.y, o));
    if (lane == 0) acc[e] = zadd(acc[e], s);
}

__global__ void zero2_kernel(double2 *p, size_t n)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = make_double2(0.0, 0.0);
}

// ---------------------------------------------------------------------------
// host side

// bump allocator over the handle's adjoint work space
struct arena {
    char *base;
    size_t off = 0;
    template <typename T> T *take(size_t n)
    {
        T *p = reinterpret_cast<T *>(base + off);
        off += (n * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

int zgemm(th_handle *h, int opA, int opB, int m, int n, int k, double2 alpha, const double2 *A, int lda,
          const double2 *B, int ldb, double2 beta, double2 *C, int ldc)
{
    if (m <= 0 || n <= 0) return TH_OK;
    dim3 grid((unsigned)((m + 15) / 16), (unsigned)((n + 15) / 16));
    hipLaunchKernelGGL(zgemm_kernel, grid, dim3(256), 0, h->stream, opA, opB, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

int zero2(th_handle *h, double2 *p, size_t n)
{
    if (n == 0) return TH_OK;
    hipLaunchKernelGGL(zero2_kernel, dim3(nblk(n, 256)), dim3(256), 0, h->stream, p, n);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

// chi0hat (and Uhat) from chihat through the Dyson step; all device pointers
int dyson_adjoint(th_handle *h, const double2 *d_chi0, const double2 *d_chi, const double2 *d_chat,
                  const double *U_host, double2 *d_c0hat, double2 *d_Uhat, arena &ar)
{
    const int D = h->D;
    const size_t DD = (size_t)D * D;
    const double2 *U = reinterpret_cast<const double2 *>(U_host);
    std::vector<char> on(D, 0);
    for (int c = 0; c < D; ++c)
        for (int a = 0; a < D; ++a) {
            double2 u = U[a + (size_t)D * c];
            if (u.x != 0.0 || u.y != 0.0) { on[a] = 1; on[c] = 1; }
        }
    std::vector<int> S;
    for (int a = 0; a < D; ++a)
        if (on[a]) S.push_back(a);
    const int r = (int)S.size();
    const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0), mone = make_double2(-1.0, 0.0);
    TH_HIP(hipMemcpyAsync(d_c0hat, d_chat, sizeof(double2) * DD, hipMemcpyDeviceToDevice, h->stream));
    if (r == 0) {
        if (d_Uhat) {
            // Uhat = -chi0^H chihat chi^H with chi = chi0
            double2 *R = ar.take<double2>(DD);
            int rc = zgemm(h, 0, 1, D, D, D, one, d_chat, D, d_chi, D, zero, R, D);
            if (!rc) rc = zgemm(h, 1, 0, D, D, D, mone, d_chi0, D, R, D, zero, d_Uhat, D);
            return rc;
        }
        return TH_OK;
    }
    std::vector<double2> US((size_t)r * r);
    for (int b = 0; b < r; ++b)
        for (int a = 0; a < r; ++a) US[a + (size_t)r * b] = U[S[a] + (size_t)D * S[b]];
    int *d_S = ar.take<int>(r);
    int *d_perm = ar.take<int>(r);
    double2 *d_US = ar.take<double2>((size_t)r * r);
    double2 *d_W = ar.take<double2>((size_t)r * r);
    double2 *c0S = ar.take<double2>((size_t)D * r);   // chi0[:,S]
    double2 *chS = ar.take<double2>((size_t)D * r);   // chi[S,:]
    double2 *Y = ar.take<double2>((size_t)D * r);
    double2 *Y2 = ar.take<double2>((size_t)D * r);
    double2 *Q = ar.take<double2>((size_t)D * r);
    TH_HIP(hipMemcpyAsync(d_S, S.data(), sizeof(int) * r, hipMemcpyHostToDevice, h->stream));
    TH_HIP(hipMemcpyAsync(d_US, US.data(), sizeof(double2) * r * r, hipMemcpyHostToDevice, h->stream));
    const unsigned g = nblk((size_t)D * r, 256);
    hipLaunchKernelGGL(gather_cols, dim3(g), dim3(256), 0, h->stream, d_chi0, D, d_S, r, c0S);
    hipLaunchKernelGGL(gather_rows, dim3(g), dim3(256), 0, h->stream, d_chi, D, d_S, r, chS);
    hipLaunchKernelGGL(form_w_adj, dim3(nblk((size_t)r * r, 256)), dim3(256), 0, h->stream, d_chi0, D, d_S, d_US, r,
                       d_W);
    hipLaunchKernelGGL(gj_inverse, dim3(1), dim3(1024), sizeof(double2) * 2 * r, h->stream, d_W, r, d_perm,
                       th_status_word(h));
    TH_HIP(hipGetLastError());
    int rc;
    // X = M^-H chihat = chihat - P W^-H U_S^H chi0[:,S]^H chihat   (X in d_c0hat)
    if ((rc = zgemm(h, 1, 0, r, D, D, one, c0S, D, d_chat, D, zero, Y, r))) return rc;
    if ((rc = zgemm(h, 1, 0, r, D, r, one, d_US, r, Y, r, zero, Y2, r))) return rc;
    if ((rc = zgemm(h, 1, 0, r, D, r, one, d_W, r, Y2, r, zero, Y, r))) return rc;
    hipLaunchKernelGGL(scatter_sub_rows, dim3(g), dim3(256), 0, h->stream, d_c0hat, D, d_S, r, Y);
    TH_HIP(hipGetLastError());
    // Q = X chi[S,:]^H  (D x r)
    if ((rc = zgemm(h, 0, 1, D, r, D, one, d_c0hat, D, chS, r, zero, Q, D))) return rc;
    if (d_Uhat) {
        // Uhat = -chi0^H X chi^H
        double2 *R = ar.take<double2>(DD);
        if ((rc = zgemm(h, 0, 1, D, D, D, one, d_c0hat, D, d_chi, D, zero, R, D))) return rc;
        if ((rc = zgemm(h, 1, 0, D, D, D, mone, d_chi0, D, R, D, zero, d_Uhat, D))) return rc;
    }
    // chi0hat = X N^H = X - (X chi[S,:]^H) U_S^H in the columns S
    if ((rc = zgemm(h, 0, 1, D, r, r, one, Q, D, d_US, r, zero, Y2, D))) return rc;
    hipLaunchKernelGGL(scatter_sub_cols, dim3(g), dim3(256), 0, h->stream, d_c0hat, D, d_S, r, Y2);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

bool bad(const void *p) { return p == nullptr; }

}  // namespace

extern "C" int th_chi_adjoint(th_handle *h, const double *q, double omega, double Ef, double eta, double etap,
                              int n_x, const double *x2, const double *p2, int precision, const double *chi0,
                              const double *U, const double *chi, const double *chi_hat, double *t_hat,
                              double *h0_hat, double *U_hat)
{
    if (!h || bad(q) || bad(chi0) || bad(chi_hat) || bad(t_hat) || bad(h0_hat) || (U && bad(chi)) ||
        (U_hat && !U) || (precision != TH_PREC_FP64 && precision != TH_PREC_MIXED) || n_x < 0 ||
        (std::fabs(omega) >= 1e-15 && n_x > 0 && (bad(x2) || bad(p2)))) {
        th_set_error("th_chi_adjoint: bad arguments");
        return TH_E_ARG;
    }
    if (!h->d_kim || h->n_im <= 0) {
        th_set_error("th_set_imag_points must be called before th_chi_adjoint");
        return TH_E_STATE;
    }
    const bool real_axis = std::fabs(omega) >= 1e-15 && n_x > 0;
    if (real_axis && (!h->d_kre || h->nk_re <= 0)) {
        th_set_error("th_set_real_kmesh must be called before th_chi_adjoint at omega != 0");
        return TH_E_STATE;
    }
    const int D = h->D, d = h->d, nO = h->nOrb;
    const size_t DD = (size_t)D * D, dd = (size_t)d * d;
    const int nout = h->nhop * nO * nO + (int)dd;
    constexpr int CHUNK = 32;     // segments per partial sum of the hop and h0 reductions
    // pairs per batch: adjoint buffers within twice the panel budget
    const size_t per = (size_t)2 * h->nA2 * 2 * h->Mpad * sizeof(double) + 5 * dd * sizeof(double2) +
                       (size_t)nout * sizeof(double2) / CHUNK + (size_t)h->nhop * sizeof(double2) +
                       (size_t)h->nA2 * nO * nO * sizeof(double2) + 64;
    const long n_re = real_axis ? (long)n_x * h->nk_re : 0;
    const long total = std::max(h->n_im, n_re);
    long bp = (long)std::max<size_t>(TH_KC, ((size_t)h->batch_mb << 21) / per);
    bp = std::min(bp, total);
    const long nch_max = (bp + CHUNK - 1) / CHUNK;
    // work space
    size_t need = 0;
    need += 4 * ((DD * sizeof(double2) + 255) & ~(size_t)255);                  // chat, c0hat, chi0, chi
    need += U_hat ? ((DD * sizeof(double2) + 255) & ~(size_t)255) * 2 : 0;       // Uhat, R
    {
        size_t r_max = (size_t)D;                                               // support size bound
        need += 2 * ((r_max * sizeof(int) + 255) & ~(size_t)255) + 2 * ((r_max * r_max * sizeof(double2) + 255) & ~(size_t)255) +
                5 * (((size_t)D * r_max * sizeof(double2) + 255) & ~(size_t)255);
    }
    need += ((size_t)nout * sizeof(double2) + 255) & ~(size_t)255;              // gradient accumulator
    need += (((size_t)nch_max * nout * sizeof(double2)) + 255) & ~(size_t)255;  // partial sums
    need += 2 * (((size_t)bp * h->nA2 * 2 * h->Mpad * sizeof(double) + 255) & ~(size_t)255);
    need += 5 * (((size_t)bp * dd * sizeof(double2) + 255) & ~(size_t)255);
    need += ((size_t)bp * std::max(1, h->nhop) * sizeof(double2) + 255) & ~(size_t)255;
    need += ((size_t)bp * h->nA2 * nO * nO * sizeof(double2) + 255) & ~(size_t)255;
    need += ((size_t)3 * (h->nhop + h->nA2) * sizeof(int) + 255) & ~(size_t)255;   // hop groups
    need += 2 * (((size_t)(h->n_im + n_re + 2) * sizeof(int) + 255) & ~(size_t)255);   // segments, their order
    need += 2 * ((((size_t)bp / TH_RUN + 1) * sizeof(float) + 255) & ~(size_t)255);   // A, B panel maxima per workgroup
    need += 4096;
    if (need > h->adj_bytes) {
        hipFree(h->d_adj);
        h->d_adj = nullptr;
        h->adj_bytes = 0;
        TH_HIP(hipMalloc(&h->d_adj, need));
        h->adj_bytes = need;
    }
    arena ar{reinterpret_cast<char *>(h->d_adj)};
    double2 *d_chat = ar.take<double2>(DD);
    double2 *d_c0hat = ar.take<double2>(DD);
    double2 *d_x0 = ar.take<double2>(DD);
    double2 *d_x = ar.take<double2>(DD);
    double2 *d_grad = ar.take<double2>(nout);
    double2 *d_part = ar.take<double2>((size_t)nch_max * nout);
    double *hatA = ar.take<double>((size_t)bp * h->nA2 * 2 * h->Mpad);
    double *hatB = ar.take<double>((size_t)bp * h->nA2 * 2 * h->Mpad);
    double2 *rawA = ar.take<double2>((size_t)bp * dd);
    double2 *rawB = ar.take<double2>((size_t)bp * dd);
    double2 *gA = ar.take<double2>((size_t)bp * dd);
    double2 *gB = ar.take<double2>((size_t)bp * dd);
    double2 *tmp = ar.take<double2>((size_t)bp * dd);
    double2 *d_ph = ar.take<double2>((size_t)bp * std::max(1, h->nhop));
    double2 *d_S = ar.take<double2>((size_t)bp * h->nA2 * nO * nO);
    double2 *d_Uhat = U_hat ? ar.take<double2>(DD) : nullptr;
    int *d_seg = ar.take<int>((size_t)(h->n_im + n_re + 2));
    int *d_sord = ar.take<int>((size_t)(h->n_im + n_re + 2));
    float *d_bmax = ar.take<float>((size_t)bp / TH_RUN + 1);
    float *d_bmaxB = ar.take<float>((size_t)bp / TH_RUN + 1);

    // segments: runs of consecutive pairs on one k-point within a batch (the
    // stored imaginary list keeps a k-point's pairs together in runs; real
    // pairs are taken grouped k-major, th_real_grouped: runs of up to TH_RUN
    // nodes of one k-point); their first pairs (indices of the whole list)
    // and the end of the last, axis 0 then axis 1; built once per pair lists,
    // batch size and node count (a pass over every pair on the host)
    std::vector<int>(&segv)[2] = h->adj_seg;
    const long skey[3] = {h->list_version, bp, real_axis ? (long)n_x : 0};
    for (int axis = 0; axis < 2 && !std::equal(skey, skey + 3, h->adj_seg_key); ++axis) {
        const long npairs = axis == 0 ? h->n_im : n_re;
        std::vector<int> &sv = segv[axis];
        sv.clear();
        if (axis == 1 && !real_axis) break;
        sv.reserve((size_t)npairs / 8 + 2);
        unsigned ik_prev = 0;
        for (long p = 0; p < npairs; ++p) {
            bool brk = p % bp == 0;
            if (axis == 0) {
                brk = brk || h->kid_im_local[p] != h->kid_im_local[p - 1];
            } else {
                unsigned ie, ik;
                th_real_grouped((unsigned)p, (unsigned)h->nk_re, (unsigned)n_x, ie, ik);
                brk = brk || ik != ik_prev;
                ik_prev = ik;
            }
            if (brk) sv.push_back((int)p);
        }
        sv.push_back((int)npairs);
    }
    if (!std::equal(skey, skey + 3, h->adj_seg_key)) {
        // rounds: per batch, a segment goes to the round after the last one
        // holding a segment of its k-point (unique k-point index)
        std::vector<int> cnt((size_t)h->n_kuniq, 0), kb, rr;
        for (int axis = 0; axis < 2; ++axis) {
            h->adj_sord[axis].clear();
            h->adj_roff[axis].clear();
            if (axis == 1 && !real_axis) break;
            const long npairs = axis == 0 ? h->n_im : n_re;
            const std::vector<int> &sv = segv[axis];
            std::vector<int> &so = h->adj_sord[axis];
            size_t b0 = 0;
            for (long p0 = 0; p0 < npairs; p0 += bp) {
                const long np = std::min(bp, npairs - p0);
                size_t b1 = b0;
                while (sv[b1] < p0 + np) ++b1;
                const int ns = (int)(b1 - b0);
                kb.assign(ns, 0);
                rr.assign(ns, 0);
                int nr = 0;
                for (int i = 0; i < ns; ++i) {
                    const long pf = sv[b0 + i];
                    if (axis == 0) {
                        kb[i] = h->kid_im_host[pf];
                    } else {
                        unsigned ie, ik;
                        th_real_grouped((unsigned)pf, (unsigned)h->nk_re, (unsigned)n_x, ie, ik);
                        kb[i] = h->kid_re_host[ik];
                    }
                    rr[i] = cnt[kb[i]]++;
                    nr = std::max(nr, rr[i] + 1);
                }
                for (int i = 0; i < ns; ++i) cnt[kb[i]] = 0;
                std::vector<long> off(nr + 1, 0);
                for (int i = 0; i < ns; ++i) ++off[rr[i] + 1];
                for (int r = 0; r < nr; ++r) off[r + 1] += off[r];
                std::vector<long> fill(off.begin(), off.end() - 1);
                const size_t base = so.size();
                so.resize(base + ns);
                for (int i = 0; i < ns; ++i) so[base + fill[rr[i]]++] = i;
                for (long &o : off) o += (long)base;
                h->adj_roff[axis].push_back(std::move(off));
                b0 = b1;
            }
        }
    }
    std::copy(skey, skey + 3, h->adj_seg_key);
    const size_t seg1 = segv[0].size();
    TH_HIP(hipMemcpyAsync(d_seg, segv[0].data(), sizeof(int) * seg1, hipMemcpyHostToDevice, h->stream));
    TH_HIP(hipMemcpyAsync(d_sord, h->adj_sord[0].data(), sizeof(int) * h->adj_sord[0].size(), hipMemcpyHostToDevice,
                          h->stream));
    if (real_axis) {
        TH_HIP(hipMemcpyAsync(d_seg + seg1, segv[1].data(), sizeof(int) * segv[1].size(), hipMemcpyHostToDevice,
                              h->stream));
        TH_HIP(hipMemcpyAsync(d_sord + seg1, h->adj_sord[1].data(), sizeof(int) * h->adj_sord[1].size(),
                              hipMemcpyHostToDevice, h->stream));
    }

    // hop groups: at most HG hops of one atom pair (stored hops are grouped by pair)
    std::vector<int> hoff(h->nA2 + 1), grp;
    TH_HIP(hipMemcpy(hoff.data(), h->d_hoff, sizeof(int) * (h->nA2 + 1), hipMemcpyDeviceToHost));
    for (int b = 0; b < h->nA2; ++b)
        for (int n = hoff[b]; n < hoff[b + 1]; n += HG) {
            grp.push_back(b);
            grp.push_back(n);
            grp.push_back(std::min(HG, hoff[b + 1] - n));
        }
    const int ngroups = (int)grp.size() / 3;
    int *d_groups = ar.take<int>(std::max<size_t>(grp.size(), 1));
    stage_timer tm(h);
    const int idall = tm.begin(4);
    for (auto &c : h->n_launch) c = 0;
    int rc = th_reset_status(h);
    if (rc) return rc;
    if (ngroups > 0)
        TH_HIP(hipMemcpyAsync(d_groups, grp.data(), sizeof(int) * grp.size(), hipMemcpyHostToDevice, h->stream));
    if ((rc = th_ensure_panels(h, bp))) return rc;
    if (real_axis && (rc = th_upload_nodes(h, n_x, x2, p2))) return rc;
    if ((rc = th_ensure_eig(h, q))) return rc;
    TH_HIP(hipMemcpyAsync(d_chat, chi_hat, sizeof(double2) * DD, hipMemcpyHostToDevice, h->stream));
    TH_HIP(hipMemcpyAsync(d_x0, chi0, sizeof(double2) * DD, hipMemcpyHostToDevice, h->stream));
    if (U) TH_HIP(hipMemcpyAsync(d_x, chi, sizeof(double2) * DD, hipMemcpyHostToDevice, h->stream));
    // 1. Dyson
    int id = tm.begin(3);
    if (U) {
        if ((rc = dyson_adjoint(h, d_x0, d_x, d_chat, U, d_c0hat, d_Uhat, ar))) return rc;
    } else {
        TH_HIP(hipMemcpyAsync(d_c0hat, d_chat, sizeof(double2) * DD, hipMemcpyDeviceToDevice, h->stream));
    }
    tm.end(id);
    // 2. accumulator cotangents (axis 0 imaginary, axis 1 real) in h->d_acc
    id = tm.begin(2);
    const size_t nacc = (size_t)2 * h->nA2 * 2 * h->Mpad * h->Mpad;
    if ((rc = th_launch_zero(h, h->d_acc, nacc))) return rc;
    hipLaunchKernelGGL(unfinalize_1, dim3(nblk(DD, 256)), dim3(256), 0, h->stream, d_c0hat, h->d_acc, h->nA, nO,
                       h->Mpad, D, real_axis ? 1 : 0);
    hipLaunchKernelGGL(unfinalize_2, dim3(nblk(DD, 256)), dim3(256), 0, h->stream, d_c0hat, h->d_acc, h->nA, nO,
                       h->Mpad, D);
    TH_HIP(hipGetLastError());
    if ((rc = zero2(h, d_grad, nout))) return rc;
    tm.end(id);
    const size_t accsz = (size_t)h->nA2 * 2 * h->Mpad * h->Mpad;
    // the B panels are recomputed like the A panels unless "adjoint_bcache" asks
    // for the forward's omega-independent cache (same E_F, eta', precision)
    const bool use_bcache = h->adj_bcache && h->d_bcache && h->bcache_prec == precision && h->bcache_key[0] == Ef &&
                            h->bcache_key[1] == etap;
    // split words in the cache are read back as (hi + lo) 2^-e
    const int split_e = (use_bcache && h->bcache_presplit) ? h->bcache_pexp : INT_MIN;
    const int ntile = (d + 31) / 32;
    // the hop and h0 cotangents of n segment (or k-point) sums Hs, whose
    // phases are in d_ph: spin sums, chunk partials, then added to d_grad
    auto hop_h0_reduce = [&](const double2 *Hs, int n) -> int {
        hreduce_params hp{};
        hp.np = n; hp.nk_re = h->nk_re;
        hp.chunk = CHUNK; hp.nout = nout; hp.nhop = h->nhop; hp.nOrb = nO; hp.d = d;
        hp.q0 = q[0]; hp.q1 = q[1]; hp.q2 = q[2];
        hp.kim = h->d_kim; hp.kre = h->d_kre; hp.R = h->d_R; hp.jb = h->d_jb; hp.ib = h->d_ib;
        hp.Hh = Hs;
        hp.nA = h->nA;
        hp.S = d_S;
        hp.ph = d_ph;
        hp.partial = d_part;
        hipLaunchKernelGGL(spin_sum_kernel, dim3(nblk((size_t)n * h->nA2 * nO * nO, 256)), dim3(256), 0, h->stream,
                           (long)n, h->nA, nO, Hs, d_S);
        const int nch = (n + CHUNK - 1) / CHUNK;
        if (ngroups > 0)
            hipLaunchKernelGGL(hop_reduce_kernel, dim3(nblk((size_t)ngroups * nO * nO, 64), nch), dim3(256), 0,
                               h->stream, hp, (const int *)d_groups, ngroups);
        hipLaunchKernelGGL(h0_reduce_kernel, dim3(nblk(dd, 64), nch), dim3(256), 0, h->stream, hp);
        hipLaunchKernelGGL(hsum_kernel, dim3(nblk(nout, 4)), dim3(256), 0, h->stream, (const double2 *)d_part, nch,
                           nout, d_grad);
        TH_HIP(hipGetLastError());
        return TH_OK;
    };
    // Hhat in the spectral form when both sides' eigenpairs are ready (then
    // neither side's G is formed: the Green stage writes panels only)
    bool spec_h = (d == 18 || d == 36 || d == 54) && th_side_eig(h, true, q) && th_side_eig(h, false, q);
    // spectral form: per side and unique k-point the eigenbasis sum W_k (FP64),
    // and per batch its segments in rounds of distinct k-points (launch after
    // launch, so that the sums add in the segments' order without races)
    const size_t wk_side = (size_t)h->n_kuniq * dd;
    std::vector<int>(&sord)[2] = h->adj_sord;
    std::vector<std::vector<long>>(&roff)[2] = h->adj_roff;
    if (spec_h) {
        const size_t wb = 2 * wk_side * sizeof(double2);
        if (h->wk_bytes < wb) {
            hipFree(h->d_wk);
            h->d_wk = nullptr;
            h->wk_bytes = 0;
            if (hipMalloc(&h->d_wk, wb) == hipSuccess) {
                h->wk_bytes = wb;
            } else {   // no room for the k-points' sums: each pair's G instead
                (void)hipGetLastError();
                h->d_wk = nullptr;
                spec_h = false;
            }
        }
        if (spec_h && (rc = th_launch_zero(h, reinterpret_cast<double *>(h->d_wk), 4 * wk_side))) return rc;
    }
    // 3. pair batches of each axis
    for (int axis = 0; axis < (real_axis ? 2 : 1); ++axis) {
        const long npairs = axis == 0 ? h->n_im : n_re;
        const double *That = h->d_acc + (size_t)axis * accsz;
        const std::vector<int> &sv = segv[axis];
        const int *dsv = d_seg + (axis == 0 ? 0 : seg1);
        if ((rc = th_panel_adjoint_prepare(h, precision, That))) return rc;
        size_t sb0 = 0;   // first segment of the batch
        int bi = 0;       // the batch's index
        for (long p0 = 0; p0 < npairs; p0 += bp, ++bi) {
            const long np = std::min(bp, npairs - p0);
            size_t sb1 = sb0;
            while (sv[sb1] < p0 + np) ++sb1;   // segments never straddle a batch
            const int nseg = (int)(sb1 - sb0);
            const int *dseg = dsv + sb0;
            th_green_args a{};
            a.p0 = p0; a.np = np; a.q[0] = q[0]; a.q[1] = q[1]; a.q[2] = q[2];
            a.omega = omega; a.Ef = Ef; a.eta = eta; a.etap = etap; a.prec = precision;
            if (axis == 1) {   // real pairs k-major: a k-point's nodes in one segment
                a.kmajor_re = 1;
                a.nx = n_x;
            }
            id = tm.begin(0);
            a.mode = axis == 0 ? TH_G_IMAG_A : TH_G_REAL_A;
            a.panel = h->d_pA;
            a.raw = spec_h ? nullptr : rawA;
            a.blkmax = precision != TH_PREC_FP64 ? d_bmax : nullptr;   // the A panel's scale for the fp16 GEMM
            if ((rc = th_launch_green(h, a))) return rc;
            a.blkmax = nullptr;
            const long nbmax = h->last_blkmax;
            // the imaginary-axis B side does not depend on omega: when the forward's
            // panel cache holds it (same E_F, eta', precision), read G from there
            const void *pB = h->d_pB;
            long nbmaxB = 0;   // workgroup maxima of a recomputed B panel
            th_rowmap brows;   // rows of pB: this batch's pairs in order
            if (axis == 0 && use_bcache) {
                // the cache holds the whole list, per batch of the forward's
                // contraction in its summation order
                pB = h->d_bcache;
                brows.pg0 = p0;
                brows.bpc = h->bcache_bp;
                brows.ntot = h->n_im;
                if (!spec_h) {   // G of the B side for the per-pair Hhat products
                    const unsigned g = nblk((size_t)np * dd, 256);
                    const float unsplit = split_e == INT_MIN ? 0.0f : ldexpf(1.0f, -split_e);
                    if (precision == TH_PREC_FP64)
                        hipLaunchKernelGGL(raw_from_bpanel<double>, dim3(g), dim3(256), 0, h->stream,
                                           (const double *)pB, np, h->nA, h->bs, h->Mpad, brows, 0.0f, rawB);
                    else
                        hipLaunchKernelGGL(raw_from_bpanel<float>, dim3(g), dim3(256), 0, h->stream,
                                           (const float *)pB, np, h->nA, h->bs, h->Mpad, brows, unsplit, rawB);
                    TH_HIP(hipGetLastError());
                }
            } else {
                a.mode = axis == 0 ? TH_G_IMAG_B : TH_G_REAL_B;
                a.panel = h->d_pB;
                a.raw = spec_h ? nullptr : rawB;
                a.blkmax = precision != TH_PREC_FP64 ? d_bmaxB : nullptr;   // the B panel's scale
                if ((rc = th_launch_green(h, a))) return rc;
                a.blkmax = nullptr;
                nbmaxB = h->last_blkmax;
            }
            tm.end(id);
            id = tm.begin(1);
            if ((rc = th_launch_panel_adjoint(h, pB, precision, That, np, 1, hatA, brows,
                                              axis == 0 && use_bcache ? split_e : INT_MIN, d_bmaxB, nbmaxB)))
                return rc;
            if ((rc = th_launch_panel_adjoint(h, h->d_pA, precision, That, np, 0, hatB, th_rowmap(), INT_MIN, d_bmax,
                                              nbmax)))
                return rc;
            h->n_launch[1] += 2;
            tm.end(id);
            id = tm.begin(2);
            for (int side = 0; side < 2; ++side) {
                assemble_params ap{};
                ap.mode = axis == 0 ? (side == 0 ? TH_G_IMAG_A : TH_G_IMAG_B) : (side == 0 ? TH_G_REAL_A : TH_G_REAL_B);
                ap.nA = h->nA; ap.bs = h->bs; ap.d = d; ap.Mpad = h->Mpad; ap.nA2 = h->nA2;
                ap.p0 = p0; ap.np = np; ap.nk_re = h->nk_re; ap.nx = n_x;
                ap.wim = h->d_wim; ap.p2 = h->d_p2; ap.wre = h->d_wre;
                ap.hat = side == 0 ? hatA : hatB;
                double2 *gh = side == 0 ? gA : gB;
                const double2 *G = side == 0 ? rawA : rawB;
                ap.out = gh;
                // Hhat = G^H Ghat G^H per pair, summed over each segment
                const double2 *Hseg = gh;
                if (spec_h) {
                    // spectral form: V and lambda of the segment's k-point, no G; the
                    // segment sums go to the k-points' eigenbasis sums (reduced after
                    // the last batch)
                    const th_eig_set *es = th_side_eig(h, side == 0, q);
                    hspec_params sp{};
                    sp.V = es->V; sp.L = es->L;
                    sp.kid = axis == 0 ? h->d_kid_im : h->d_kid_re;
                    sp.yim = h->d_yim; sp.x2 = h->d_x2;
                    // z = E_F + omega + i (y + eta), E_F + i (y + eta'), x + omega + i eta, x + i eta'
                    sp.er0 = (axis == 0 ? Ef : 0.0) + (side == 0 ? omega : 0.0);
                    sp.ei0 = side == 0 ? eta : etap;
                    const bool f32 = precision != TH_PREC_FP64;
                    const int nblk3 = (d / 3) * (d / 3);
                    int ppw = std::max(1, 512 / nblk3);
                    while (ppw > 1 && hspec_lds(d, ppw, f32) > 64 * 1024) --ppw;
                    if (d == 18) ppw = std::min(ppw, 4);   // segments of 16 pairs: four pairs per thread slot
                    const int nt = ((ppw * nblk3 + 63) / 64) * 64;
                    const size_t lds = hspec_lds(d, ppw, f32);
                    double2 *Wk = h->d_wk + (size_t)side * wk_side;
                    const std::vector<long> &off = roff[axis][bi];
                    const int *dso = d_sord + (axis == 0 ? 0 : seg1);
                    for (size_t r = 0; r + 1 < off.size(); ++r) {
                        const unsigned nw = (unsigned)(off[r + 1] - off[r]);
                        auto launch = [&](auto kern) {
                            TH_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                            hipLaunchKernelGGL(kern, dim3(nw), dim3(nt), lds, h->stream, ap, sp, dseg, dso + off[r], ppw,
                                               Wk);
                            return 0;
                        };
                        if (d == 18) rc = f32 ? launch(hhat_spec_kernel<float, 18>) : launch(hhat_spec_kernel<double, 18>);
                        else if (d == 36) rc = f32 ? launch(hhat_spec_kernel<float, 36>) : launch(hhat_spec_kernel<double, 36>);
                        else rc = f32 ? launch(hhat_spec_kernel<float, 54>) : launch(hhat_spec_kernel<double, 54>);
                        if (rc) return rc;
                        TH_HIP(hipGetLastError());
                    }
                    continue;
                } else if (d <= 54 && d % 3 == 0) {
                    // the pairs' G, Ghat in LDS, 3 x 3 register blocks, one workgroup per segment
                    const int nblk3 = (d / 3) * (d / 3);
                    // at most ~42 KB of LDS (three workgroups per CU); FP32 products
                    // for the mixed path
                    const bool f32 = precision != TH_PREC_FP64;
                    const size_t esz = f32 ? sizeof(float2) : sizeof(double2);
                    const int ppw = std::max(1, std::min(42 * 1024 / (int)(2 * dd * esz), 320 / nblk3));
                    const int nt = ((ppw * nblk3 + 63) / 64) * 64;
                    const size_t lds = std::max((size_t)ppw * 2 * dd * esz, (size_t)ppw * dd * sizeof(double2));
                    auto launch = [&](auto kern) {
                        TH_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                        hipLaunchKernelGGL(kern, dim3((unsigned)nseg), dim3(nt), lds, h->stream, ap, G, dseg, ppw, gh);
                        return 0;
                    };
                    if (d == 18) rc = f32 ? launch(hhat_seg_d_kernel<float, 18>) : launch(hhat_seg_d_kernel<double, 18>);
                    else if (d == 36) rc = f32 ? launch(hhat_seg_d_kernel<float, 36>) : launch(hhat_seg_d_kernel<double, 36>);
                    else if (d == 54) rc = f32 ? launch(hhat_seg_d_kernel<float, 54>) : launch(hhat_seg_d_kernel<double, 54>);
                    else rc = f32 ? launch(hhat_seg_kernel<float>) : launch(hhat_seg_kernel<double>);
                    if (rc) return rc;
                    TH_HIP(hipGetLastError());
                } else {
                    hipLaunchKernelGGL(assemble_kernel, dim3(nblk((size_t)np * dd, 256)), dim3(256), 0, h->stream, ap);
                    TH_HIP(hipGetLastError());
                    // Hhat = G^H (Ghat G^H)
                    const unsigned nwg = (unsigned)(np * ntile * ntile);
                    hipLaunchKernelGGL(bmm_kernel, dim3(nwg), dim3(256), 0, h->stream, (const double2 *)gh, 0, G, 1, tmp,
                                       d, ntile);
                    hipLaunchKernelGGL(bmm_kernel, dim3(nwg), dim3(256), 0, h->stream, G, 1, (const double2 *)tmp, 0, gh,
                                       d, ntile);
                    hipLaunchKernelGGL(seg_sum_kernel, dim3(nblk((size_t)nseg * dd, 256)), dim3(256), 0, h->stream,
                                       (const double2 *)gh, dseg, p0, nseg, (int)dd, tmp);
                    TH_HIP(hipGetLastError());
                    Hseg = tmp;
                }
                hipLaunchKernelGGL(phase_kernel, dim3(nblk((size_t)nseg * std::max(1, h->nhop), 256)), dim3(256), 0,
                                   h->stream, dseg, nseg, h->nhop, axis == 0 ? 1 : 0, side == 0 ? 1 : 0, (int)h->nk_re,
                                   n_x, q[0], q[1], q[2], (const double *)h->d_kim, (const double *)h->d_kre,
                                   (const double *)h->d_R, d_ph);
                TH_HIP(hipGetLastError());
                if ((rc = hop_h0_reduce(Hseg, nseg))) return rc;
                TH_HIP(hipGetLastError());
            }
            tm.end(id);
            sb0 = sb1;
        }
    }
    if (spec_h) {
        // Hhat_k = V_k W_k V_k^H per side and unique k-point, then the phase,
        // spin-sum and hop/h0 reductions over the k-points, bp at a time
        id = tm.begin(2);
        for (int side = 0; side < 2; ++side) {
            const th_eig_set *es = th_side_eig(h, side == 0, q);
            const double2 *Wk = h->d_wk + (size_t)side * wk_side;
            for (long k0 = 0; k0 < h->n_kuniq; k0 += bp) {
                const int nkc = (int)std::min(bp, h->n_kuniq - k0);
                const size_t lds = 2 * dd * sizeof(double2);
                auto fin = [&](auto kern) {
                    TH_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                    hipLaunchKernelGGL(kern, dim3((unsigned)nkc), dim3(256), lds, h->stream, Wk, es->V, k0, gA);
                    return 0;
                };
                rc = d == 18 ? fin(hhat_final_kernel<18>) : d == 36 ? fin(hhat_final_kernel<36>) : fin(hhat_final_kernel<54>);
                if (rc) return rc;
                hipLaunchKernelGGL(phase_k_kernel, dim3(nblk((size_t)nkc * std::max(1, h->nhop), 256)), dim3(256), 0,
                                   h->stream, (const double *)h->d_kuniq, k0, nkc, h->nhop, side == 0 ? 1 : 0, q[0], q[1],
                                   q[2], (const double *)h->d_R, d_ph);
                TH_HIP(hipGetLastError());
                if ((rc = hop_h0_reduce(gA, nkc))) return rc;
            }
        }
        tm.end(id);
    }
    tm.end(idall);
    // 4. results to the host, hops back in the caller's order
    std::vector<double2> g(nout);
    TH_HIP(hipMemcpyAsync(g.data(), d_grad, sizeof(double2) * nout, hipMemcpyDeviceToHost, h->stream));
    if (U_hat) TH_HIP(hipMemcpyAsync(U_hat, d_Uhat, sizeof(double2) * DD, hipMemcpyDeviceToHost, h->stream));
    TH_HIP(hipStreamSynchronize(h->stream));
    const int nO2 = nO * nO;
    double2 *th = reinterpret_cast<double2 *>(t_hat);
    for (int n = 0; n < h->nhop; ++n)
        for (int e = 0; e < nO2; ++e) th[(size_t)h->hop_order[n] * nO2 + e] = g[(size_t)n * nO2 + e];
    double2 *h0o = reinterpret_cast<double2 *>(h0_hat);
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) h0o[r + (size_t)d * c] = g[(size_t)h->nhop * nO2 + (size_t)r * d + c];
    rc = th_check_status(h);
    tm.collect();
    return rc;
}
