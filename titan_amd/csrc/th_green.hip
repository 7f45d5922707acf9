// th_green.hip -- batched Green functions G(k,z) = (z - H(k))^-1 for gfx950.
//
// One workgroup per (k,E) pair.  The workgroup assembles H(k) from h0 and the
// hop table in LDS (mod_hamiltonian.F90:35-87 and 183-243), inverts (z - H)
// in place by Gauss-Jordan elimination with partial (row) pivoting
// (the result of LAPACK zgetrf+zgetri as called by invers,
// mod_tools.F90:796-829), and scatters the atom blocks the contraction needs
// into a GEMM panel (the gf(:,:,i,j) scatter of mod_greenfunction.F90:55-62,
// plus the weight and the real-axis combinations of eintshechi.F90:143-165).
//
// Precision: the elimination always runs in FP64 (on CDNA4 unpacked FP32
// VALU work issues at the FP64 rate, so an FP32 elimination would not be
// faster); the mixed mode rounds the finished G once to FP32 when it writes
// the panel.  For the adjoint the kernel can also write the unpermuted G
// (FP64, row-major d x d per pair) to a second buffer.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "th_internal.h"

namespace {

template <typename R> struct cplx { R x, y; };

template <typename R> __device__ inline cplx<R> cmul(cplx<R> a, cplx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R> __device__ inline cplx<R> csub(cplx<R> a, cplx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> __device__ inline cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> __device__ inline cplx<R> conj(cplx<R> a) { return {a.x, -a.y}; }
// r x + z w (r real)
__device__ inline cplx<double> rcfma(double r, cplx<double> x, cplx<double> z, cplx<double> w)
{
    return {fma(r, x.x, fma(z.x, w.x, -z.y * w.y)), fma(r, x.y, fma(z.x, w.y, z.y * w.x))};
}
template <typename R> __device__ inline cplx<R> cinv(cplx<R> a)
{
    R n = a.x * a.x + a.y * a.y;
    return {a.x / n, -a.y / n};
}
template <typename R> __device__ inline R cabs1(cplx<R> a) { return fabs(a.x) + fabs(a.y); }

struct green_params {
    int mode, prec;
    int nA, nOrb, bs, d, nhop, Mpad, nA2;
    long p0, np, nk_re;
    int nx, kmajor_re;       // real axis: nx nodes; pairs grouped k-major when kmajor_re (th_real_grouped), else p = ie * nk_re + ik
    double q0, q1, q2, omega, Ef, eta, etap;
    const double *omega_p;   // when set, omega is read from device memory (graph replays)
    const double *R;
    const int *jb, *ib;
    const double2 *t, *h0;
    const int *hoff;
    const double *kim, *yim, *wim, *kre, *wre, *x2, *p2;
    void *panel;                 // panel rows of the batch
    long slot0;                  // batch index of this launch's first pair
    unsigned nrun_out;           // summation order: see panel_row
    int presplit, pexp, pnx;     // fp16 split panels (th_green_args::presplit)
    double pareal;
    double2 *raw;
    int *status;
    // spectral mode: eigenpairs of H(k) per k-point of the current sets and
    // the k-point index of every pair (imaginary list, real mesh)
    const double2 *eigV;
    const double *eigL;
    const _Float16 *eigS;        // the split planes of eigV (th_eig_set::S), when kept
    const int *kid_im, *kid_re;
    float *blkmax;               // th_green_args::blkmax
};

// Panel row of batch pair b = slot0 + pl.  With nrun_out = n > 0 the first
// TH_RUN n pairs (n runs of TH_RUN pairs of one k-point each) are stored
// transposed, pair j of run r in row j n + r, so that consecutive rows, the
// order in which the contraction sums them, belong to different k-points
// (sums over runs of one k-point round coherently in FP32); the rest keep
// their order.
__device__ inline size_t panel_row(const green_params &g, long pl)
{
    const unsigned b = (unsigned)(g.slot0 + pl), n = g.nrun_out;
    return (n != 0u && b < (unsigned)TH_RUN * n) ? (size_t)((b % TH_RUN) * n + b / TH_RUN) : (size_t)b;
}

// In-place Gauss-Jordan inverse of the d x d row-major matrix M in LDS.
// Row interchanges are recorded; the final column permutation is returned
// in idx (G[r][c] = M[r][idx[c]]).  Returns through *status the 1-based
// step of the first exactly-zero pivot (0 if none).
template <typename R, int NT>
__device__ void gj_invert(cplx<R> *M, int d, cplx<R> *colk, cplx<R> *prow, int *perm, int *idx,
                          int *s_piv, cplx<R> *s_inv, int *status)
{
    const int tid = threadIdx.x;
    for (int k = 0; k < d; ++k) {
        // pivot: largest |re|+|im| in column k at or below the diagonal,
        // first index on ties (izamax rule); rows < 64 live in wave 0
        if (tid < 64) {
            R v = (tid >= k && tid < d) ? cabs1(M[tid * d + k]) : R(-1);
            int ix = tid;
            for (int off = 32; off > 0; off >>= 1) {
                R ov = __shfl_xor(v, off);
                int oi = __shfl_xor(ix, off);
                if (ov > v || (ov == v && oi < ix)) { v = ov; ix = oi; }
            }
            if (tid == 0) { *s_piv = ix; perm[k] = ix; }
        }
        __syncthreads();
        const int p = *s_piv;
        cplx<R> rk, rp, ck, piv;
        if (tid < d) {
            rk = M[k * d + tid];
            rp = M[p * d + tid];
            ck = (tid == k) ? M[p * d + k] : (tid == p ? M[k * d + k] : M[tid * d + k]);
            piv = M[p * d + k];
        }
        __syncthreads();
        if (tid < d) {
            if (piv.x == R(0) && piv.y == R(0)) {
                if (tid == 0) atomicMin(status, k + 1);
                piv.x = R(1);
            }
            cplx<R> inv = cinv(piv);
            if (p != k) M[p * d + tid] = rk;
            cplx<R> pr = (tid == k) ? inv : cmul(rp, inv);
            M[k * d + tid] = pr;
            prow[tid] = pr;
            colk[tid] = ck;
            if (tid == 0) *s_inv = inv;
        }
        __syncthreads();
        const cplx<R> inv = *s_inv;
        for (int e = tid; e < d * d; e += NT) {
            int r = e / d, c = e - r * d;
            if (r == k) continue;
            cplx<R> f = colk[r];
            if (c == k) {
                cplx<R> v = cmul(f, inv);
                M[e] = {-v.x, -v.y};
            } else {
                M[e] = csub(M[e], cmul(f, prow[c]));
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int c = 0; c < d; ++c) idx[c] = c;
        for (int k = d - 1; k >= 0; --k) {
            int pk = perm[k];
            if (pk != k) { int tmp = idx[k]; idx[k] = idx[pk]; idx[pk] = tmp; }
        }
    }
    __syncthreads();
}


// (k, z, weight) of pair p for the kernel mode
// node ie and k-point ik of real-axis pair p (pair counts below 2^31,
// checked on the host, so the divisions are 32-bit).  Node-major:
// p = ie * nk_re + ik; grouped k-major (kmajor_re): th_real_grouped.
__device__ inline void real_split(const green_params &g, long p, long &ie, long &ik)
{
    // both layouts computed, the result selected by value (no branch-dependent
    // stores through the references, which the compiler would keep in memory)
    const unsigned up = (unsigned)p, nk = (unsigned)g.nk_re;
    const unsigned a0 = up / nk, ie0 = a0, ik0 = up - a0 * nk;
    unsigned ie1, ik1;
    th_real_grouped(up, nk, g.nx > 0 ? (unsigned)g.nx : 1u, ie1, ik1);
    ie = (long)(g.kmajor_re ? ie1 : ie0);
    ik = (long)(g.kmajor_re ? ik1 : ik0);
}

__device__ inline void pair_point(const green_params &g, long p, double &k0, double &k1, double &k2, double &er,
                                  double &ei, double &w)
{
    w = 1.0;
    const double omega = g.omega_p ? *g.omega_p : g.omega;
    if (g.mode == TH_G_IMAG_A || g.mode == TH_G_IMAG_B) {
        k0 = g.kim[3 * p]; k1 = g.kim[3 * p + 1]; k2 = g.kim[3 * p + 2];
        if (g.mode == TH_G_IMAG_A) {
            k0 += g.q0; k1 += g.q1; k2 += g.q2;
            er = g.Ef + omega; ei = g.yim[p] + g.eta; w = g.wim[p];
        } else {
            er = g.Ef; ei = g.yim[p] + g.etap;
        }
    } else {
        long ie, ik;
        real_split(g, p, ie, ik);
        k0 = g.kre[3 * ik]; k1 = g.kre[3 * ik + 1]; k2 = g.kre[3 * ik + 2];
        if (g.mode == TH_G_REAL_A) {
            k0 += g.q0; k1 += g.q1; k2 += g.q2;
            er = g.x2[ie] + omega; ei = g.eta; w = g.p2[ie] * g.wre[ik];
        } else {
            er = g.x2[ie]; ei = g.etap;
        }
    }
}

// Writes the panel row of pair pl from the inverse held row-major in S with
// column permutation idx (G[r][c] = S[r][idx[c]]).
template <typename P, int NT>
__device__ void write_panel(const green_params &g, long pl, const cplx<double> *S, const int *idx, int d, double w)
{
    const int tid = threadIdx.x;
    const int bs = g.bs, Mb = bs * bs, Mpad = g.Mpad, nA = g.nA;
    P *out = reinterpret_cast<P *>(g.panel) + panel_row(g, pl) * g.nA2 * 2 * Mpad;
    for (int ij = 0; ij < g.nA2; ++ij) {
        const int i = ij / nA, j = ij - (ij / nA) * nA;
        P *o = out + (size_t)ij * 2 * Mpad;
        for (int e = tid; e < Mpad; e += NT) {
            double vr = 0.0, vi = 0.0;
            if (e < Mb) {
                int rr = e / bs, cc = e - rr * bs;
                if (g.mode == TH_G_IMAG_A) {          // w * Gq(i,j)
                    cplx<double> v = S[(i * bs + rr) * d + idx[j * bs + cc]];
                    vr = w * v.x; vi = w * v.y;
                } else if (g.mode == TH_G_IMAG_B) {   // G(j,i)
                    cplx<double> v = S[(j * bs + rr) * d + idx[i * bs + cc]];
                    vr = v.x; vi = v.y;
                } else if (g.mode == TH_G_REAL_A) {   // -i w (Gq(i,j) - Gq^dagger(i,j))
                    cplx<double> a = S[(i * bs + rr) * d + idx[j * bs + cc]];
                    cplx<double> b = S[(j * bs + cc) * d + idx[i * bs + rr]];
                    double dr = a.x - b.x, di = a.y + b.y;      // a - conj(b)
                    vr = w * di; vi = -w * dr;                  // -i w (dr + i di)
                } else {                                        // G^dagger(j,i) = conj(G(i,j))^T
                    cplx<double> v = S[(i * bs + cc) * d + idx[j * bs + rr]];
                    vr = v.x; vi = -v.y;
                }
            }
            o[e] = (P)vr;
            o[Mpad + e] = (P)vi;
        }
    }
}

// Register-resident Gauss-Jordan for a compile-time d = DD.  Thread t owns
// the entries e = t + i*NT (row-major) of (z - H) in registers; per
// elimination step only the pivot column and the two swapped rows travel
// through LDS (double-buffered by step parity, two barriers per step).
// The pivot search is repeated by every wave on the published column, so
// no third barrier is needed.
template <typename P, int DD, int NT>
__global__ __launch_bounds__(NT) void green_reg_kernel(green_params g)
{
    constexpr int NE = DD * DD;
    constexpr int EPT = (NE + NT - 1) / NT;
    constexpr bool FULL_STAGE = (NE * 16) <= 49152;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double> *S = reinterpret_cast<cplx<double> *>(smem);          // staging of the result
    cplx<double> *ph = S + (FULL_STAGE ? NE : 2 * g.bs * g.bs);        // nhop phases
    __shared__ cplx<double> colk[2][DD], rowk[2][DD], rowp[2][DD];
    __shared__ int perm[DD], idx[DD];
    const int tid = threadIdx.x, lane = tid & 63;
    const long pl = blockIdx.x;
    const long p = g.p0 + pl;
    double k0, k1, k2, er, ei, w;
    pair_point(g, p, k0, k1, k2, er, ei, w);
    for (int n = tid; n < g.nhop; n += NT) {
        double a = k0 * g.R[3 * n] + k1 * g.R[3 * n + 1] + k2 * g.R[3 * n + 2];
        double sn, cs;
        sincos(a, &sn, &cs);
        ph[n] = {cs, sn};
    }
    __syncthreads();
    // owned entries of z - H(k)
    const int nO = g.nOrb, nO2 = nO * nO;
    // entry u of this thread: e = tid + u*NT, row e / DD, column e % DD
#define RR(u) ((tid + (u) * NT) < NE ? (tid + (u) * NT) / DD : -1)
#define CC(u) ((tid + (u) * NT) % DD)
    cplx<double> v[EPT];
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
        const int e = tid + u * NT;
        v[u] = {0.0, 0.0};
        if (e < NE) {
            const int r = e / DD, c = e - (e / DD) * DD;
            double2 hv = g.h0[r + (size_t)c * DD];
            cplx<double> x = {(r == c ? er : 0.0) - hv.x, (r == c ? ei : 0.0) - hv.y};
            const int ja = r / (2 * nO), rs = (r / nO) & 1, nu = r - (r / nO) * nO;
            const int ia = c / (2 * nO), cs = (c / nO) & 1, mu = c - (c / nO) * nO;
            if (rs == cs) {
                const int h1 = g.hoff[ja * g.nA + ia], h2 = g.hoff[ja * g.nA + ia + 1];
                for (int n = h1; n < h2; ++n) {
                    double2 tv = g.t[(size_t)n * nO2 + nu + nO * mu];
                    x = csub(x, cmul(cplx<double>{tv.x, tv.y}, ph[n]));
                }
            }
            v[u] = x;
        }
    }
    for (int k = 0; k < DD; ++k) {
        const int bb = k & 1;
        // publish column k
#pragma unroll
        for (int u = 0; u < EPT; ++u)
            if (RR(u) >= 0 && CC(u) == k) colk[bb][RR(u)] = v[u];
        __syncthreads();
        // pivot search (every wave): largest |re|+|im| at or below the diagonal
        double pv = -1.0;
        int pi = DD;
        for (int r = lane; r < DD; r += 64)
            if (r >= k) {
                double a = cabs1(colk[bb][r]);
                if (a > pv) { pv = a; pi = r; }
            }
        for (int off = 32; off > 0; off >>= 1) {
            double ov = __shfl_xor(pv, off);
            int oi = __shfl_xor(pi, off);
            if (ov > pv || (ov == pv && oi < pi)) { pv = ov; pi = oi; }
        }
        const int pr = pi;
        if (tid == 0) {
            perm[k] = pr;
            if (pv == 0.0) atomicMin(g.status, k + 1);
        }
        // publish rows k and pr
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            if (RR(u) == k) rowk[bb][CC(u)] = v[u];
            if (RR(u) == pr) rowp[bb][CC(u)] = v[u];
        }
        __syncthreads();
        cplx<double> piv = colk[bb][pr];
        if (piv.x == 0.0 && piv.y == 0.0) piv.x = 1.0;
        const cplx<double> inv = cinv(piv);
        const cplx<double> fk = colk[bb][k];   // multiplier of the row that moves to pr
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int r = RR(u), c = CC(u);
            if (r < 0) continue;
            if (r == k) {
                v[u] = (c == k) ? inv : cmul(rowp[bb][c], inv);
            } else {
                const cplx<double> src = (r == pr) ? rowk[bb][c] : v[u];
                const cplx<double> f = (r == pr) ? fk : colk[bb][r];
                if (c == k) {
                    cplx<double> t = cmul(f, inv);
                    v[u] = {-t.x, -t.y};
                } else {
                    v[u] = csub(src, cmul(f, cmul(rowp[bb][c], inv)));
                }
            }
        }
    }
    if (tid == 0) {
        for (int c = 0; c < DD; ++c) idx[c] = c;
        for (int k = DD - 1; k >= 0; --k) {
            int pk = perm[k];
            if (pk != k) { int tmp = idx[k]; idx[k] = idx[pk]; idx[pk] = tmp; }
        }
    }
    if constexpr (FULL_STAGE) {
#pragma unroll
        for (int u = 0; u < EPT; ++u)
            if (RR(u) >= 0) S[RR(u) * DD + CC(u)] = v[u];
        __syncthreads();
        write_panel<P, NT>(g, pl, S, idx, DD, w);
    } else {
        // the inverse does not fit LDS: stage one atom-block pair at a time.
        // Entry (r, c') of the registers is G[r][pos[c']] with pos = idx^-1.
        int *pos = perm;   // perm is no longer needed
        __syncthreads();
        for (int r = tid; r < DD; r += NT) pos[idx[r]] = r;
        __syncthreads();
        const int bs = g.bs, Mb = bs * bs, Mpad = g.Mpad, nA = g.nA;
        cplx<double> *Sij = S, *Sji = S + Mb;          // blocks (i,j) and (j,i), row-major bs x bs
        P *out = reinterpret_cast<P *>(g.panel) + panel_row(g, pl) * g.nA2 * 2 * Mpad;
        for (int ij = 0; ij < g.nA2; ++ij) {
            const int i = ij / nA, j = ij - (ij / nA) * nA;
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                const int r = RR(u);
                if (r < 0) continue;
                const int fcu = pos[CC(u)];
                const int ra = r / bs, ca = fcu / bs;
                if (ra == i && ca == j) Sij[(r - i * bs) * bs + fcu - j * bs] = v[u];
                if (ra == j && ca == i) Sji[(r - j * bs) * bs + fcu - i * bs] = v[u];
            }
            __syncthreads();
            P *o = out + (size_t)ij * 2 * Mpad;
            for (int e = tid; e < Mpad; e += NT) {
                double vr = 0.0, vi = 0.0;
                if (e < Mb) {
                    const int a = e / bs, b = e - a * bs;
                    if (g.mode == TH_G_IMAG_A) {
                        cplx<double> x = Sij[a * bs + b];
                        vr = w * x.x; vi = w * x.y;
                    } else if (g.mode == TH_G_IMAG_B) {       // G(j,i)
                        cplx<double> x = Sji[a * bs + b];
                        vr = x.x; vi = x.y;
                    } else if (g.mode == TH_G_REAL_A) {       // -i w (G(i,j) - G(j,i)^dagger)
                        cplx<double> x = Sij[a * bs + b], y = Sji[b * bs + a];
                        double dr = x.x - y.x, di = x.y + y.y;
                        vr = w * di; vi = -w * dr;
                    } else {                                  // G^dagger(j,i) = conj(G(i,j))^T
                        cplx<double> x = Sij[b * bs + a];
                        vr = x.x; vi = -x.y;
                    }
                }
                o[e] = (P)vr;
                o[Mpad + e] = (P)vi;
            }
            __syncthreads();
        }
    }
}
#undef RR
#undef CC


// 2-D register-blocked Gauss-Jordan for compile-time d = DD.  Thread (ty,tx)
// of a TY x TX grid owns rows ty + i*TY and columns tx + j*TX of (z - H).
// Pivoting is implicit: rows are never moved; step k picks the unused row
// p_k with the largest |re|+|im| in column k (one packed-key DPP wavefront
// max, repeated by every wave) and the result is unscrambled at the end,
// G[k][p_c] = V[p_k][c].  The published scaled pivot row carries 1 + 1/piv
// in column k, so every non-pivot entry is updated by one complex FMA,
// v -= f * row, with no selects.  Per step: the pivot column and the scaled
// pivot row pass through LDS (double-buffered by step parity), two barriers.
extern "C" __device__ unsigned __ockl_wfred_max_u32(unsigned);

// Gauss-Jordan elimination with implicit pivoting on a register tile: thread
// (ty,tx) of a TY x TX grid owns rows ty + i*TY and columns tx + j*TX of the
// DD x DD matrix in v.  Rows are never moved; step k picks the unused row
// p_k with the largest |re|+|im| in column k (one packed-key wavefront max,
// repeated by every wave).  On return v holds V with G[k][p_c] = V[p_k][c]:
// perm[k] = p_k and qinv[p_k] = k.  'used' must be zero on entry.
template <int DD, int TY, int TX>
__device__ __forceinline__ void gj2d_eliminate(cplx<double> (&v)[(DD + TY - 1) / TY][(DD + TX - 1) / TX],
                                               cplx<double> (*colk)[DD], cplx<double> (*prs)[DD], int *perm,
                                               int *qinv, unsigned char *used, int *status)
{
    constexpr int NT = TY * TX;
    constexpr int RI = (DD + TY - 1) / TY, CJ = (DD + TX - 1) / TX;
    static_assert(DD <= 255, "row index packed in 8 bits");
    static_assert(CJ <= 8, "the column loop is unrolled over the CJ register slots");
    const int tid = threadIdx.x, lane = tid & 63;
    const int ty = tid / TX, tx = tid - (tid / TX) * TX;
    // Column k = kx + jb*TX lives in register slot jb of the lanes with
    // tx == kx: with jb unrolled, the pivot column is published from a static
    // register slot.  The pivot row is data dependent (ip = pr / TY is uniform
    // across the workgroup, so the tests on it are scalar branches).
#pragma unroll
    for (int jb = 0; jb < CJ; ++jb) {
        for (int kx = 0; kx < TX; ++kx) {
            const int k = kx + jb * TX;
            if (k >= DD) break;
            const int bb = k & 1;
            if (tx == kx) {
#pragma unroll
                for (int i = 0; i < RI; ++i) {
                    const int r = ty + i * TY;
                    if (r < DD) colk[bb][r] = v[i][jb];
                }
            }
            __syncthreads();
            // pivot: key = magnitude (float, low 8 mantissa bits dropped) | 255 - row,
            // so the maximum key is the largest |re|+|im|, lowest row on ties
            unsigned key = 0u;
            for (int r = lane; r < DD; r += 64)
                if (!used[r]) {
                    const float a = (float)cabs1(colk[bb][r]);
                    const unsigned kk = (__float_as_uint(a) & 0xFFFFFF00u) | (unsigned)(255 - r);
                    key = kk > key ? kk : key;
                }
            key = __ockl_wfred_max_u32(key);
            const int pr = 255 - (int)(key & 0xFFu);
            const int ip = pr / TY, typ = pr - ip * TY;
            cplx<double> piv = colk[bb][pr];
            if (piv.x == 0.0 && piv.y == 0.0) {
                if (tid == 0) atomicMin(status, k + 1);
                piv.x = 1.0;
            }
            const cplx<double> inv = cinv(piv);
            // the owners of row pr publish it scaled by 1/piv, with 1 + 1/piv in
            // column k, so every other row is updated by v -= f * prs
#pragma unroll
            for (int i = 0; i < RI; ++i)
                if (i == ip && ty == typ) {
#pragma unroll
                    for (int j = 0; j < CJ; ++j) {
                        const int c = tx + j * TX;
                        if (c < DD) prs[bb][c] = (c == k) ? cplx<double>{1.0 + inv.x, inv.y} : cmul(v[i][j], inv);
                    }
                }
            __syncthreads();
            // after the barrier: every wave has finished this step's pivot search
            if (tid == 0) {
                perm[k] = pr;
                used[pr] = 1;
            }
            cplx<double> pj[CJ];
#pragma unroll
            for (int j = 0; j < CJ; ++j) {
                const int c = tx + j * TX;
                pj[j] = (c < DD) ? prs[bb][c] : cplx<double>{0.0, 0.0};
            }
#pragma unroll
            for (int i = 0; i < RI; ++i) {
                const int r = ty + i * TY;
                const cplx<double> f = (r < DD) ? colk[bb][r] : cplx<double>{0.0, 0.0};
#pragma unroll
                for (int j = 0; j < CJ; ++j) {
                    // v -= f * pj
                    v[i][j] = {fma(-f.x, pj[j].x, fma(f.y, pj[j].y, v[i][j].x)),
                               fma(-f.x, pj[j].y, fma(-f.y, pj[j].x, v[i][j].y))};
                }
            }
            // row pr takes the published values: v/piv, and 1/piv in column k
#pragma unroll
            for (int i = 0; i < RI; ++i)
                if (i == ip && ty == typ) {
#pragma unroll
                    for (int j = 0; j < CJ; ++j) {
                        const int c = tx + j * TX;
                        v[i][j] = (c == k) ? inv : pj[j];
                    }
                }
        }
    }
    // unscramble: G[k][p_c] = V[p_k][c], i.e. entry (rho, c) goes to (q[rho], perm[c])
    __syncthreads();
    for (int r = tid; r < DD; r += NT) qinv[perm[r]] = r;   // NT may be below DD
    __syncthreads();
}

// The whole inverse is staged through LDS for the panel scatter when it
// takes at most 12 KB per wave of the workgroup (so the staging buffer does
// not limit occupancy below what the registers allow); otherwise one pair of
// atom blocks at a time.
constexpr bool full_stage(int dd, int nt) { return dd * dd * 16 <= 12288 * (nt / 64); }

template <typename P, int DD, int TY, int TX>
__global__ __launch_bounds__(TY *TX) void green_2d_kernel(green_params g)
{
    constexpr int NT = TY * TX;
    constexpr int RI = (DD + TY - 1) / TY, CJ = (DD + TX - 1) / TX;
    constexpr int NE = DD * DD;
    constexpr bool FULL_STAGE = full_stage(DD, NT);
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double> *S = reinterpret_cast<cplx<double> *>(smem);
    cplx<double> *ph = S + (FULL_STAGE ? NE : 2 * g.bs * g.bs);
    __shared__ cplx<double> colk[2][DD], prs[2][DD];
    __shared__ int perm[DD], qinv[DD], ident[DD];
    __shared__ unsigned char used[DD];
    const int tid = threadIdx.x, lane = tid & 63;
    const int ty = tid / TX, tx = tid - (tid / TX) * TX;
    const long pl = blockIdx.x;
    const long p = g.p0 + pl;
    double k0, k1, k2, er, ei, w;
    pair_point(g, p, k0, k1, k2, er, ei, w);
    for (int n = tid; n < g.nhop; n += NT) {
        double a = k0 * g.R[3 * n] + k1 * g.R[3 * n + 1] + k2 * g.R[3 * n + 2];
        double sn, cs;
        sincos(a, &sn, &cs);
        ph[n] = {cs, sn};
    }
    for (int r = tid; r < DD; r += NT) {
        used[r] = 0;
        ident[r] = r;
    }
    __syncthreads();
    const int nO = g.nOrb, nO2 = nO * nO;
    cplx<double> v[RI][CJ];
#pragma unroll
    for (int i = 0; i < RI; ++i)
#pragma unroll
        for (int j = 0; j < CJ; ++j) {
            const int r = ty + i * TY, c = tx + j * TX;
            v[i][j] = {0.0, 0.0};
            if (r < DD && c < DD) {
                double2 hv = g.h0[r + (size_t)c * DD];
                cplx<double> x = {(r == c ? er : 0.0) - hv.x, (r == c ? ei : 0.0) - hv.y};
                const int ja = r / (2 * nO), rs = (r / nO) & 1, nu = r - (r / nO) * nO;
                const int ia = c / (2 * nO), csp = (c / nO) & 1, mu = c - (c / nO) * nO;
                if (rs == csp) {
                    const int h1 = g.hoff[ja * g.nA + ia], h2 = g.hoff[ja * g.nA + ia + 1];
                    for (int n = h1; n < h2; ++n) {
                        double2 tv = g.t[(size_t)n * nO2 + nu + nO * mu];
                        x = csub(x, cmul(cplx<double>{tv.x, tv.y}, ph[n]));
                    }
                }
                v[i][j] = x;
            }
        }
    gj2d_eliminate<DD, TY, TX>(v, colk, prs, perm, qinv, used, g.status);
    int dcol[CJ];
#pragma unroll
    for (int j = 0; j < CJ; ++j) dcol[j] = (tx + j * TX < DD) ? perm[tx + j * TX] : -1;
    if constexpr (FULL_STAGE) {
#pragma unroll
        for (int i = 0; i < RI; ++i) {
            const int r = ty + i * TY;
            if (r >= DD) continue;
            const int dr = qinv[r];
#pragma unroll
            for (int j = 0; j < CJ; ++j)
                if (dcol[j] >= 0) S[dr * DD + dcol[j]] = v[i][j];
        }
        __syncthreads();
        if (g.raw) {
            double2 *o = g.raw + (size_t)pl * NE;
            for (int e = tid; e < NE; e += NT) o[e] = make_double2(S[e].x, S[e].y);
        }
        write_panel<P, NT>(g, pl, S, ident, DD, w);
    } else {
        const int bs = g.bs, Mb = bs * bs, Mpad = g.Mpad, nA = g.nA;
        cplx<double> *Sij = S, *Sji = S + Mb;
        P *out = reinterpret_cast<P *>(g.panel) + panel_row(g, pl) * g.nA2 * 2 * Mpad;
        int drow[RI];
#pragma unroll
        for (int i = 0; i < RI; ++i) drow[i] = (ty + i * TY < DD) ? qinv[ty + i * TY] : -1;
        if (g.raw) {
            double2 *o = g.raw + (size_t)pl * NE;
#pragma unroll
            for (int i = 0; i < RI; ++i) {
                if (drow[i] < 0) continue;
#pragma unroll
                for (int j = 0; j < CJ; ++j)
                    if (dcol[j] >= 0) o[drow[i] * DD + dcol[j]] = make_double2(v[i][j].x, v[i][j].y);
            }
        }
        for (int ij = 0; ij < g.nA2; ++ij) {
            const int i0 = ij / nA, j0 = ij - (ij / nA) * nA;
#pragma unroll
            for (int i = 0; i < RI; ++i) {
                if (drow[i] < 0) continue;
                const int ra = drow[i] / bs;
#pragma unroll
                for (int j = 0; j < CJ; ++j) {
                    if (dcol[j] < 0) continue;
                    const int ca = dcol[j] / bs;
                    if (ra == i0 && ca == j0) Sij[(drow[i] - i0 * bs) * bs + dcol[j] - j0 * bs] = v[i][j];
                    if (ra == j0 && ca == i0) Sji[(drow[i] - j0 * bs) * bs + dcol[j] - i0 * bs] = v[i][j];
                }
            }
            __syncthreads();
            P *o = out + (size_t)ij * 2 * Mpad;
            for (int e = tid; e < Mpad; e += NT) {
                double vr = 0.0, vi = 0.0;
                if (e < Mb) {
                    const int a = e / bs, b = e - a * bs;
                    if (g.mode == TH_G_IMAG_A) {
                        cplx<double> x = Sij[a * bs + b];
                        vr = w * x.x; vi = w * x.y;
                    } else if (g.mode == TH_G_IMAG_B) {
                        cplx<double> x = Sji[a * bs + b];
                        vr = x.x; vi = x.y;
                    } else if (g.mode == TH_G_REAL_A) {
                        cplx<double> x = Sij[a * bs + b], y = Sji[b * bs + a];
                        double dr = x.x - y.x, di = x.y + y.y;
                        vr = w * di; vi = -w * dr;
                    } else {
                        cplx<double> x = Sij[b * bs + a];
                        vr = x.x; vi = -x.y;
                    }
                }
                o[e] = (P)vr;
                o[Mpad + e] = (P)vi;
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// d = 2b beyond what one CU holds (d = 288: 1.3 MB per matrix): 2 x 2 block
// elimination.  With (z - H) = [[A, B], [C, E]] in b x b blocks and H
// Hermitian, Im z > 0, both A and the Schur complement S = E - C A^-1 B are
// invertible and their inverses have norm at most 1/Im z (they are
// compressions of the resolvent), so no pivoting across blocks is needed:
//   G = [[A^-1 - X G_BA, -X S^-1], [-S^-1 Y, S^-1]],  X = A^-1 B, Y = C A^-1.
// The two b x b inversions use the register-resident elimination above;
// the six b x b products run on a batched complex GEMM.  Matrices are
// row-major in global memory, one per pair.

// M = z - H(k), row-major d x d per pair
__global__ __launch_bounds__(256) void assemble_m_kernel(green_params g, double2 *M)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double> *ph = reinterpret_cast<cplx<double> *>(smem);
    const int tid = threadIdx.x, d = g.d, nO = g.nOrb, nO2 = nO * nO;
    const long pl = blockIdx.x;
    double k0, k1, k2, er, ei, w;
    pair_point(g, g.p0 + pl, k0, k1, k2, er, ei, w);
    for (int n = tid; n < g.nhop; n += blockDim.x) {
        double a = k0 * g.R[3 * n] + k1 * g.R[3 * n + 1] + k2 * g.R[3 * n + 2];
        double sn, cs;
        sincos(a, &sn, &cs);
        ph[n] = {cs, sn};
    }
    __syncthreads();
    double2 *o = M + (size_t)pl * d * d;
    for (int e = tid; e < d * d; e += blockDim.x) {
        const int r = e / d, c = e - (e / d) * d;
        double2 hv = g.h0[r + (size_t)c * d];
        cplx<double> x = {(r == c ? er : 0.0) - hv.x, (r == c ? ei : 0.0) - hv.y};
        const int ja = r / (2 * nO), rs = (r / nO) & 1, nu = r - (r / nO) * nO;
        const int ia = c / (2 * nO), csp = (c / nO) & 1, mu = c - (c / nO) * nO;
        if (rs == csp) {
            const int h1 = g.hoff[ja * g.nA + ia], h2 = g.hoff[ja * g.nA + ia + 1];
            for (int n = h1; n < h2; ++n) {
                double2 tv = g.t[(size_t)n * nO2 + nu + nO * mu];
                x = csub(x, cmul(cplx<double>{tv.x, tv.y}, ph[n]));
            }
        }
        o[e] = make_double2(x.x, x.y);
    }
}

// out = in^-1 for one b x b block per workgroup (row-major, leading
// dimensions and pair strides given)
template <int DD, int TY, int TX>
__global__ __launch_bounds__(TY *TX) void inv_block_kernel(const double2 *in, long in_stride, int in_ld,
                                                           double2 *out, long out_stride, int out_ld, int *status)
{
    constexpr int NT = TY * TX;
    constexpr int RI = (DD + TY - 1) / TY, CJ = (DD + TX - 1) / TX;
    __shared__ cplx<double> colk[2][DD], prs[2][DD];
    __shared__ int perm[DD], qinv[DD];
    __shared__ unsigned char used[DD];
    const int tid = threadIdx.x;
    const int ty = tid / TX, tx = tid - (tid / TX) * TX;
    const long pl = blockIdx.x;
    for (int r = tid; r < DD; r += NT) used[r] = 0;
    __syncthreads();
    const double2 *src = in + pl * in_stride;
    cplx<double> v[RI][CJ];
#pragma unroll
    for (int i = 0; i < RI; ++i)
#pragma unroll
        for (int j = 0; j < CJ; ++j) {
            const int r = ty + i * TY, c = tx + j * TX;
            v[i][j] = {0.0, 0.0};
            if (r < DD && c < DD) {
                const double2 x = src[(size_t)r * in_ld + c];
                v[i][j] = {x.x, x.y};
            }
        }
    gj2d_eliminate<DD, TY, TX>(v, colk, prs, perm, qinv, used, status);
    double2 *dst = out + pl * out_stride;
#pragma unroll
    for (int i = 0; i < RI; ++i) {
        const int r = ty + i * TY;
        if (r >= DD) continue;
        const int dr = qinv[r];
#pragma unroll
        for (int j = 0; j < CJ; ++j) {
            const int c = tx + j * TX;
            if (c < DD) dst[(size_t)dr * out_ld + perm[c]] = make_double2(v[i][j].x, v[i][j].y);
        }
    }
}

// Out = alpha A B + beta Cin per pair (row-major, complex FP64, real alpha
// and beta); 32 x 32 output tile per 256-thread workgroup, K staged by 16.
__global__ __launch_bounds__(256) void zgemm_batched_kernel(int m, int n, int k, double alpha, const double2 *A,
                                                            long sa, int lda, const double2 *B, long sb, int ldb,
                                                            double beta, const double2 *Cin, long sc, int ldc,
                                                            double2 *Out, long so, int ldo, int ntn)
{
    __shared__ cplx<double> As[32][17], Bs[16][33];
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const long pl = blockIdx.y;
    const int i0 = (blockIdx.x / ntn) * 32, j0 = (blockIdx.x % ntn) * 32;
    const double2 *Ap = A + pl * sa, *Bp = B + pl * sb;
    cplx<double> acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = {0.0, 0.0};
    for (int k0 = 0; k0 < k; k0 += 16) {
        for (int u = 0; u < 2; ++u) {
            const int v = t + 256 * u;
            const int row = v >> 4, kk = v & 15;
            const int gi = i0 + row, gk = k0 + kk;
            double2 x = make_double2(0.0, 0.0);
            if (gi < m && gk < k) x = Ap[(size_t)gi * lda + gk];
            As[row][kk] = {x.x, x.y};
            const int kr = v >> 5, col = v & 31;
            const int gk2 = k0 + kr, gj = j0 + col;
            double2 y = make_double2(0.0, 0.0);
            if (gk2 < k && gj < n) y = Bp[(size_t)gk2 * ldb + gj];
            Bs[kr][col] = {y.x, y.y};
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const cplx<double> a0 = As[ty][kk], a1 = As[ty + 16][kk];
            const cplx<double> b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
            acc[0][0] = {fma(a0.x, b0.x, fma(-a0.y, b0.y, acc[0][0].x)), fma(a0.x, b0.y, fma(a0.y, b0.x, acc[0][0].y))};
            acc[0][1] = {fma(a0.x, b1.x, fma(-a0.y, b1.y, acc[0][1].x)), fma(a0.x, b1.y, fma(a0.y, b1.x, acc[0][1].y))};
            acc[1][0] = {fma(a1.x, b0.x, fma(-a1.y, b0.y, acc[1][0].x)), fma(a1.x, b0.y, fma(a1.y, b0.x, acc[1][0].y))};
            acc[1][1] = {fma(a1.x, b1.x, fma(-a1.y, b1.y, acc[1][1].x)), fma(a1.x, b1.y, fma(a1.y, b1.x, acc[1][1].y))};
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const int gi = i0 + ty + 16 * a, gj = j0 + tx + 16 * b;
            if (gi >= m || gj >= n) continue;
            double re = alpha * acc[a][b].x, im = alpha * acc[a][b].y;
            if (beta != 0.0) {
                const double2 c = Cin[pl * sc + (size_t)gi * ldc + gj];
                re += beta * c.x;
                im += beta * c.y;
            }
            Out[pl * so + (size_t)gi * ldo + gj] = make_double2(re, im);
        }
}

// panels (and the optional raw copy) from G held row-major in global memory
template <typename P, int NT>
__global__ __launch_bounds__(NT) void panel_from_global_kernel(green_params g, const double2 *G)
{
    __shared__ int ident[TH_MAX_D];
    const int tid = threadIdx.x, d = g.d;
    const long pl = blockIdx.x;
    double k0, k1, k2, er, ei, w;
    pair_point(g, g.p0 + pl, k0, k1, k2, er, ei, w);
    for (int r = tid; r < d; r += NT) ident[r] = r;
    __syncthreads();
    const double2 *Gp = G + (size_t)pl * d * d;
    if (g.raw) {
        double2 *o = g.raw + (size_t)pl * d * d;
        for (int e = tid; e < d * d; e += NT) o[e] = Gp[e];
    }
    write_panel<P, NT>(g, pl, reinterpret_cast<const cplx<double> *>(Gp), ident, d, w);
}

// ---------------------------------------------------------------------------
// Spectral Green functions.  H(k) does not depend on z, so one eigen-
// decomposition H(k) = V diag(lambda) V^H per k-point serves every energy and
// frequency of a sweep, and each G(k,z) = V diag(1/(z - lambda)) V^H is one
// complex product on the MFMA units instead of an elimination.  Valid for
// Hermitian H (checked when the model is uploaded).

// Eigen-decomposition of H(k) for a list of k-points by cyclic two-sided
// Jacobi in parallel (round-robin) order, one workgroup per k-point.
// Rotation of the pair (p,q): with apq = |apq| e^{i phi} the 2 x 2 block is
// made real by diag(1, e^{-i phi}) and diagonalised by the symmetric Schur
// rotation (c, s) of tau = (aqq - app) / (2 |apq|); J = [[c, s], [-s e^{-i
// phi}, c e^{-i phi}]], A <- J^H A J, V <- V J.  Written per column m of a
// round: (M J)[:,m] = alpha_m M[:,m] + beta_m M[:,pi_m] with pi_m the partner
// of m, alpha_p = c, beta_p = -s e^{-i phi}, alpha_q = c e^{-i phi},
// beta_q = s; the rows of J^H A use the conjugate coefficients.
//
// The NP pairs of a round cover every index once, so A splits into NP x NP
// disjoint 2 x 2 blocks (rows of pair j1, columns of pair j2); each thread
// transforms whole blocks (rows, then columns, in registers), so one pass
// over A per round with one read and one write per entry.  A lives in LDS.
// V lives in registers: lane m of every wave holds column m of V for the
// rows r = wave + NW i, and the partner column comes from lane pi_m by a
// cross-lane shuffle.  Two barriers per round (coefficients, then update).
// Sweeps until the off-diagonal Frobenius norm is below 1e-15 of the whole
// (at most 30 sweeps).  Output V row-major (V[r][m] = component r of
// eigenvector m) and lambda.
template <int DD, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void eig_jacobi_kernel(green_params g, const double *klist, double q0, double q1,
                                                        double q2, double2 *Vout, double *Lout)
{
    static_assert(DD % 2 == 0 && DD <= 64, "even DD <= 64 (one lane per column)");
    constexpr int NP = DD / 2, NR = DD - 1, NW = NT / 64, RPW = (DD + NW - 1) / NW;
    constexpr int TPC = NT / NP;                  // threads per column pair in the block update
    constexpr int NA = TPC * NP;                  // threads that take part in it
    constexpr int BPT = (NP + TPC - 1) / TPC;     // row pairs per such thread
    static_assert(TPC >= 1 && NT % 64 == 0, "block shape");
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double> *ph = reinterpret_cast<cplx<double> *>(smem);
    __shared__ cplx<double> A[DD * DD];
    // rotation of column m of the pair (p, q): J[p][p] = J[q][q] = c (real),
    // J[q][p] = -conj(sigma), J[p][q] = sigma, so every column and row update is
    // c x_m + coef_m x_partner(m) with coef_p = -conj(sigma), coef_q = sigma
    __shared__ double rr[DD];
    __shared__ cplx<double> cc[DD];
    __shared__ int pt[DD];
    __shared__ double red[2 * NW];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const long kb = blockIdx.x;
    const double k0 = klist[3 * kb] + q0, k1 = klist[3 * kb + 1] + q1, k2 = klist[3 * kb + 2] + q2;
    for (int n = tid; n < g.nhop; n += NT) {
        double a = k0 * g.R[3 * n] + k1 * g.R[3 * n + 1] + k2 * g.R[3 * n + 2];
        double sn, cs;
        sincos(a, &sn, &cs);
        ph[n] = {cs, sn};
    }
    __syncthreads();
    const int nO = g.nOrb, nO2 = nO * nO;
    for (int e = tid; e < DD * DD; e += NT) {
        const int r = e / DD, c = e - (e / DD) * DD;
        double2 hv = g.h0[r + (size_t)c * DD];
        cplx<double> x = {hv.x, hv.y};
        const int ja = r / (2 * nO), rsp = (r / nO) & 1, nu = r - (r / nO) * nO;
        const int ia = c / (2 * nO), csp = (c / nO) & 1, mu = c - (c / nO) * nO;
        if (rsp == csp) {
            const int h1 = g.hoff[ja * g.nA + ia], h2 = g.hoff[ja * g.nA + ia + 1];
            for (int n = h1; n < h2; ++n) {
                double2 tv = g.t[(size_t)n * nO2 + nu + nO * mu];
                const cplx<double> tp = cmul(cplx<double>{tv.x, tv.y}, ph[n]);
                x.x += tp.x;
                x.y += tp.y;
            }
        }
        A[e] = x;
    }
    cplx<double> v[RPW];
#pragma unroll
    for (int i = 0; i < RPW; ++i) v[i] = {wave + NW * i == lane ? 1.0 : 0.0, 0.0};
    // pair j of round rnd
    auto pp = [](int j, int rnd) { return j == 0 ? NR : (rnd + j) % NR; };
    auto qq = [](int j, int rnd) { return j == 0 ? rnd : (rnd - j + NR) % NR; };
    const int j2 = tid % NP, jr0 = tid / NP;     // column pair and first row pair of this thread
    const int vm = lane < DD ? lane : 0;         // column of V held by this lane
    __syncthreads();
    for (int sweep = 0; sweep < 30; ++sweep) {
        for (int rnd = 0; rnd < NR; ++rnd) {
            if (tid < NP) {
                // zero A[p][q] = b with dd = (A[q][q] - A[p][p]) / 2, r = sqrt(dd^2 + |b|^2),
                // u = |dd| + r: c = u / sqrt(2 r u), sigma = sign(dd) b / sqrt(2 r u)
                // (t = sign(dd) |b| / u is the smaller root of t^2 + (dd / |b|) 2 t - 1)
                const int p = pp(tid, rnd), q = qq(tid, rnd);
                const cplx<double> b = A[p * DD + q];
                const double m2 = b.x * b.x + b.y * b.y;
                const double dd = 0.5 * (A[q * DD + q].x - A[p * DD + p].x);
                double c = 1.0;
                cplx<double> sg = {0.0, 0.0};
                if (m2 > 0.0) {
                    const double r = sqrt(fma(dd, dd, m2)), u = fabs(dd) + r;
                    const double kk = 1.0 / sqrt(2.0 * r * u), ks = dd >= 0.0 ? kk : -kk;
                    c = u * kk;
                    sg = {ks * b.x, ks * b.y};
                }
                rr[p] = c;
                cc[p] = {-sg.x, sg.y};
                pt[p] = q;
                rr[q] = c;
                cc[q] = sg;
                pt[q] = p;
            }
            __syncthreads();
            if (tid < NA) {
                const int p2 = pp(j2, rnd), q2 = qq(j2, rnd);
                const double c2 = rr[p2];
                const cplx<double> B2 = cc[p2], A2 = cc[q2];
#pragma unroll
                for (int b = 0; b < BPT; ++b) {
                    const int j1 = jr0 + TPC * b;
                    if (j1 < NP) {
                        const int p1 = pp(j1, rnd), q1 = qq(j1, rnd);
                        const double c1 = rr[p1];
                        const cplx<double> B1 = conj(cc[p1]), A1 = conj(cc[q1]);
                        const cplx<double> x00 = A[p1 * DD + p2], x01 = A[p1 * DD + q2];
                        const cplx<double> x10 = A[q1 * DD + p2], x11 = A[q1 * DD + q2];
                        // rows p1, q1 of J^H A at columns p2, q2
                        const cplx<double> r0p = rcfma(c1, x00, B1, x10);
                        const cplx<double> r0q = rcfma(c1, x01, B1, x11);
                        const cplx<double> r1p = rcfma(c1, x10, A1, x00);
                        const cplx<double> r1q = rcfma(c1, x11, A1, x01);
                        // then columns p2, q2 of (J^H A) J
                        A[p1 * DD + p2] = rcfma(c2, r0p, B2, r0q);
                        A[p1 * DD + q2] = rcfma(c2, r0q, A2, r0p);
                        A[q1 * DD + p2] = rcfma(c2, r1p, B2, r1q);
                        A[q1 * DD + q2] = rcfma(c2, r1q, A2, r1p);
                    }
                }
            }
            {
                // column m of V J: c v_m + coef_m v_partner(m)
                const double rm = rr[vm];
                const cplx<double> cm = cc[vm];
                const int src = lane < DD ? pt[vm] : lane;
#pragma unroll
                for (int i = 0; i < RPW; ++i) {
                    const cplx<double> y = {__shfl(v[i].x, src, 64), __shfl(v[i].y, src, 64)};
                    v[i] = rcfma(rm, v[i], cm, y);
                }
            }
            __syncthreads();
        }
        // convergence: off-diagonal against total Frobenius norm
        double off = 0.0, tot = 0.0;
        for (int e = tid; e < DD * DD; e += NT) {
            const cplx<double> x = A[e];
            const double m2 = x.x * x.x + x.y * x.y;
            tot += m2;
            if (e / DD != e - (e / DD) * DD) off += m2;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o, 64);
            tot += __shfl_xor(tot, o, 64);
        }
        if (lane == 0) { red[wave] = off; red[NW + wave] = tot; }
        __syncthreads();
        double so = 0.0, st = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) { so += red[w]; st += red[NW + w]; }
        __syncthreads();   // red is read by every thread before the next sweep writes it
        if (!(so > 1e-30 * st)) break;
    }
    double2 *vo = Vout + (size_t)kb * DD * DD;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        const int r = wave + NW * i;
        if (r < DD && lane < DD) vo[(size_t)r * DD + lane] = make_double2(v[i].x, v[i].y);
    }
    for (int m = tid; m < DD; m += NT) Lout[(size_t)kb * DD + m] = A[m * DD + m].x;
}

typedef double sdv4 __attribute__((ext_vector_type(4)));

// G(k,z) = V diag(1/(z - lambda)) V^H for one pair per workgroup on the
// FP64 MFMA units (v_mfma_f64_16x16x4: lane l holds A[l & 15][l >> 4],
// B[l >> 4][l & 15], result rows (l >> 4) + 4q, column l & 15).  Wave w owns
// the output rows 16w .. 16w+15; V of the pair's k-point is staged in LDS
// and then overwritten by G, which is scattered into the panel as by the
// elimination kernels.
// MFMA flavour of the spectral product: FP64 (v_mfma_f64_16x16x4, result
// rows (l >> 4) + 4q) or, in the mixed mode, FP32 (v_mfma_f32_16x16x4,
// result rows 4 (l >> 4) + q) with V and 1/(z - lambda) rounded to FP32.
template <typename R> struct smfma;
template <> struct smfma<double> {
    typedef sdv4 acc;
    static __device__ inline acc mac(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int g, int q) { return g + 4 * q; }
};
typedef float sfv4 __attribute__((ext_vector_type(4)));
template <> struct smfma<float> {
    typedef sfv4 acc;
    static __device__ inline acc mac(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int g, int q) { return 4 * g + q; }
};

template <typename P, int DD>
__global__ __launch_bounds__(256) void green_spectral_kernel(green_params g)
{
    typedef P R;   // arithmetic of the product: FP64 panels -> FP64, FP32 panels -> FP32
    constexpr int NB = (DD + 15) / 16;        // 16-wide blocks per dimension (waves used)
    constexpr int KS = (DD + 3) / 4;          // k-steps of 4
    __shared__ cplx<double> S[DD * DD];
    __shared__ cplx<double> f[DD];
    __shared__ int ident[DD];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const long pl = blockIdx.x;
    const long p = g.p0 + pl;
    double k0, k1, k2, er, ei, w;
    pair_point(g, p, k0, k1, k2, er, ei, w);
    const bool imag = g.mode == TH_G_IMAG_A || g.mode == TH_G_IMAG_B;
    long ie, ik;
    if (!imag) real_split(g, p, ie, ik);
    const long kb = imag ? g.kid_im[p] : g.kid_re[ik];
    const double2 *Vk = g.eigV + (size_t)kb * DD * DD;
    for (int e = tid; e < DD * DD; e += 256) {
        const double2 v = Vk[e];
        S[e] = {v.x, v.y};
    }
    for (int m = tid; m < DD; m += 256) {
        const double a = er - g.eigL[(size_t)kb * DD + m];
        const double n = a * a + ei * ei;
        f[m] = {a / n, -ei / n};
        ident[m] = m;
    }
    __syncthreads();
    typename smfma<R>::acc accr[NB], acci[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j)
        for (int q = 0; q < 4; ++q) { accr[j][q] = 0; acci[j][q] = 0; }
    if (wave < NB) {
        const int row = 16 * wave + (lane & 15);
#pragma unroll 2
        for (int ks = 0; ks < KS; ++ks) {
            const int m = 4 * ks + (lane >> 4);
            cplx<double> a = {0.0, 0.0};
            if (row < DD && m < DD) a = cmul(S[row * DD + m], f[m]);
            const R ar = (R)a.x, ai = (R)a.y, nai = -ai;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int col = 16 * j + (lane & 15);
                R br = 0, bi = 0;
                if (col < DD && m < DD) {
                    const cplx<double> v = S[col * DD + m];
                    br = (R)v.x;                           // (V^H)[m][col]
                    bi = (R)(-v.y);
                }
                accr[j] = smfma<R>::mac(ar, br, accr[j]);
                accr[j] = smfma<R>::mac(nai, bi, accr[j]);
                acci[j] = smfma<R>::mac(ar, bi, acci[j]);
                acci[j] = smfma<R>::mac(ai, br, acci[j]);
            }
        }
    }
    __syncthreads();   // every wave has read V before it is overwritten by G
    if (wave < NB) {
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 16 * wave + smfma<R>::row(lane >> 4, q), c = 16 * j + (lane & 15);
                if (r < DD && c < DD) S[r * DD + c] = {(double)accr[j][q], (double)acci[j][q]};
            }
    }
    __syncthreads();
    if (g.raw) {
        double2 *o = g.raw + (size_t)pl * DD * DD;
        for (int e = tid; e < DD * DD; e += 256) o[e] = make_double2(S[e].x, S[e].y);
    }
    write_panel<P, 256>(g, pl, S, ident, DD, w);
}

// Spectral Green functions for runs of pairs that share a k-point (the pair
// lists are ordered k-major): workgroup b takes the CH pairs b*CH ..
// b*CH+CH-1 of the batch, loads V^H of a k-point into LDS and V into
// registers once per run, and for each pair forms M = V diag(d) V^H on the
// matrix cores with the diagonal of its panel mode:
//   imag A  w f        (w Gq)
//   imag B  f          (G, stored as the (j,i) block of panel (i,j))
//   real A  2 w Im f   (-i w (Gq - Gq^H) = 2 w V Im(f) V^H, Hermitian)
//   real B  conj(f)    (G^H, stored as the (j,i) block of panel (i,j))
// f_m = 1/(z - lambda_m).  Results go from the accumulators straight into
// the panel (no staging of G), padding entries are written as zeros.
template <typename P, int DD, int CH>
__global__ __launch_bounds__(64 * ((DD + 15) / 16)) __attribute__((amdgpu_waves_per_eu(DD == 18 ? 4 : 1))) void
green_spectral_chunk_kernel(green_params g)
{
    typedef P R;
    constexpr int NB = (DD + 15) / 16, NT = 64 * NB, KS = (DD + 3) / 4, KP = 4 * KS, NC = 16 * NB;
    __shared__ cplx<R> Vh[KP * NC];      // Vh[m][c] = conj(V[c][m]); zero for m or c >= DD
    __shared__ double lam[KP];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const long s0 = (long)blockIdx.x * CH;
    const long s1 = s0 + CH < g.np ? s0 + CH : g.np;
    const bool imag = g.mode == TH_G_IMAG_A || g.mode == TH_G_IMAG_B;
    const bool bside = g.mode == TH_G_IMAG_B || g.mode == TH_G_REAL_B;
    const int bs = g.bs, nA = g.nA, Mpad = g.Mpad, Mb = bs * bs;
    for (int e = tid; e < KP * NC; e += NT) Vh[e] = {R(0), R(0)};
    // panel offsets of this lane's accumulator entries: row r = 16 wave +
    // row(q), column c = 16 j + (lane & 15); atom blocks r / bs, c / bs
    int rI[4], rR[4], cI[NB], cR[NB];
    bool rv[4], cv[NB];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 16 * wave + smfma<R>::row(lane >> 4, q);
        rv[q] = r < DD;
        rI[q] = r / bs;
        rR[q] = (r - rI[q] * bs) * bs;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int c = 16 * j + (lane & 15);
        cv[j] = c < DD;
        cI[j] = c / bs;
        cR[j] = c - cI[j] * bs;
    }
    const int arow = 16 * wave + (lane & 15);   // row of V in this lane's A operand
    long cur = -1;
    cplx<double> av[KS];
    for (long pl = s0; pl < s1; ++pl) {
        const long p = g.p0 + pl;
        double k0, k1, k2, er, ei, w;
        pair_point(g, p, k0, k1, k2, er, ei, w);
        long kb;
        if (imag) {
            kb = g.kid_im[p];
        } else {
            long ie, ik;
            real_split(g, p, ie, ik);
            kb = g.kid_re[ik];
        }
        if (kb != cur) {                        // same kb in every thread
            __syncthreads();                    // the previous run's reads of Vh are done
            const double2 *Vk = g.eigV + (size_t)kb * DD * DD;
            for (int e = tid; e < DD * DD; e += NT) {
                const int c = e / DD, m = e - (e / DD) * DD;
                const double2 x = Vk[e];
                Vh[m * NC + c] = {(R)x.x, (R)(-x.y)};
            }
            for (int m = tid; m < KP; m += NT) lam[m] = m < DD ? g.eigL[(size_t)kb * DD + m] : 0.0;
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int m = 4 * ks + (lane >> 4);
                cplx<double> v = {0.0, 0.0};
                if (arow < DD && m < DD) {
                    const double2 x = Vk[(size_t)arow * DD + m];
                    v = {x.x, x.y};
                }
                av[ks] = v;
            }
            cur = kb;
        }
        // diagonal d = (ca Re f + cb Im f) + i cc Im f for the mode
        double ca = 1.0, cb = 0.0, cc = 1.0;                  // imag B: f
        if (g.mode == TH_G_IMAG_A) { ca = w; cc = w; }        // w f
        if (g.mode == TH_G_REAL_A) { ca = 0.0; cb = 2.0 * w; cc = 0.0; }   // 2 w Im f
        if (g.mode == TH_G_REAL_B) cc = -1.0;                 // conj(f)
        typename smfma<R>::acc accr[NB], acci[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j)
            for (int q = 0; q < 4; ++q) { accr[j][q] = 0; acci[j][q] = 0; }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int m = 4 * ks + (lane >> 4);
            const double a = er - lam[m], n = a * a + ei * ei;
            const double fr = a / n, fi = -ei / n;
            const cplx<double> dg = {ca * fr + cb * fi, cc * fi};
            const cplx<double> x = cmul(av[ks], dg);
            const R ar = (R)x.x, ai = (R)x.y, nai = -ai;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const cplx<R> b = Vh[m * NC + 16 * j + (lane & 15)];
                accr[j] = smfma<R>::mac(ar, b.x, accr[j]);
                accr[j] = smfma<R>::mac(nai, b.y, accr[j]);
                acci[j] = smfma<R>::mac(ar, b.y, acci[j]);
                acci[j] = smfma<R>::mac(ai, b.x, acci[j]);
            }
        }
        P *out = reinterpret_cast<P *>(g.panel) + panel_row(g, pl) * g.nA2 * 2 * Mpad;
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (rv[q] && cv[j]) {
                    const int blk = bside ? cI[j] * nA + rI[q] : rI[q] * nA + cI[j];
                    const size_t o = (size_t)blk * 2 * Mpad + rR[q] + cR[j];
                    out[o] = (P)accr[j][q];
                    out[o + Mpad] = (P)acci[j][q];
                }
        const int npad = Mpad - Mb;
        for (int e = tid; e < g.nA2 * 2 * npad; e += NT) {
            const int part = e / npad;
            out[(size_t)part * Mpad + Mb + (e - part * npad)] = P(0);
        }
    }
}

// Spectral Green functions for d above 64, where V^H of a k-point no longer
// fits in LDS: the columns of G = V diag(d) V^H in tiles of TC, one workgroup
// per (chunk of CH pairs, column tile) and one wave per 16 rows.  V^H of the
// tile sits in LDS (reloaded when the k-point changes), the diagonal d of
// the current pair too; the A operand V[r][m] d_m is read from device memory
// per K-step (a k-point's V stays in L2 over its run of pairs).  The product
// runs on the 16x16x4 matrix cores in the panels' precision; panels as the
// run kernel writes them, or, where the caller asks for raw G, G itself (the
// launcher then forms the panels from it).
template <typename P, int DD, int TC, int RPW, int CH>
__global__ __launch_bounds__(64 * (((DD + 15) / 16 + RPW - 1) / RPW)) void green_spectral_tiled_kernel(green_params g)
{
    typedef P R;
    // NB row blocks of 16, RPW of them per wave (wave w owns blocks w, w + NW, ...)
    constexpr int NB = (DD + 15) / 16, NW = (NB + RPW - 1) / RPW, NT = 64 * NW;
    constexpr int KS = (DD + 3) / 4, KP = 4 * KS, TB = TC / 16;
    static_assert(TC % 16 == 0, "column tile of whole 16-wide blocks");
    static_assert(DD % 4 == 0 && KP == DD, "eigen-indices in whole K steps");
    __shared__ cplx<R> Vh[KP * TC];      // Vh[m][c - c0] = conj(V[c][m]); zero outside the matrix
    __shared__ cplx<double> dg[KP];      // diagonal of the current pair
    __shared__ double lam[KP];           // eigenvalues of the current run's k-point
    __shared__ double pzr[CH], pzi[CH], pzw[CH];   // Re z, Im z and weight of the workgroup's pairs
    __shared__ int pkb[CH];              // their eigen-decomposition
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int c0 = (int)blockIdx.y * TC;
    const long s0 = (long)blockIdx.x * CH;
    const long s1 = s0 + CH < g.np ? s0 + CH : g.np;
    const bool imag = g.mode == TH_G_IMAG_A || g.mode == TH_G_IMAG_B;
    const bool bside = g.mode == TH_G_IMAG_B || g.mode == TH_G_REAL_B;
    const int bs = g.bs, nA = g.nA, Mpad = g.Mpad, Mb = bs * bs;
    int rI[RPW][4], rR[RPW][4], cI[TB], cR[TB], arow[RPW];
    bool rv[RPW][4], cv[TB];
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        const int rb = wave + NW * i;
        arow[i] = 16 * rb + (lane & 15);         // row of V in this lane's A operand
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * rb + smfma<R>::row(lane >> 4, q);
            rv[i][q] = rb < NB && r < DD;
            rI[i][q] = r / bs;
            rR[i][q] = (r - rI[i][q] * bs) * bs;
        }
    }
#pragma unroll
    for (int j = 0; j < TB; ++j) {
        const int c = c0 + 16 * j + (lane & 15);
        cv[j] = c < DD;
        cI[j] = c / bs;
        cR[j] = c - cI[j] * bs;
    }
    // rows past the matrix (d = 72) read row DD - 1 and are zeroed
    int arc[RPW];
    bool arv[RPW];
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        arv[i] = arow[i] < DD;
        arc[i] = arv[i] ? arow[i] : DD - 1;
    }
    // the pairs' scalars, read once
    if (tid < CH && s0 + tid < s1) {
        const long p = g.p0 + s0 + tid;
        double k0, k1, k2, er, ei, w;
        pair_point(g, p, k0, k1, k2, er, ei, w);
        long kb;
        if (imag) {
            kb = g.kid_im[p];
        } else {
            long ie, ik;
            real_split(g, p, ie, ik);
            kb = g.kid_re[ik];
        }
        pzr[tid] = er;
        pzi[tid] = ei;
        pzw[tid] = w;
        pkb[tid] = (int)kb;
    }
    long cur = -1;
    const double2 *Vk = nullptr;
    for (long pl = s0; pl < s1; ++pl) {
        const int jj = (int)(pl - s0);
        __syncthreads();                        // the scalars; the previous pair's reads of dg, Vh, lam done
        const long kb = pkb[jj];
        const double er = pzr[jj], ei = pzi[jj], w = pzw[jj];
        if (kb != cur) {                        // same kb in every thread
            Vk = g.eigV + (size_t)kb * DD * DD;
            // consecutive threads read consecutive eigen-indices of one row of
            // V; the index is clamped (no branch per load) and the entry zeroed
            for (int e = tid; e < KP * TC; e += NT) {
                const int cc = e / KP, m = e - (e / KP) * KP, c = c0 + cc;
                const bool in = c < DD;
                const double2 v = Vk[(size_t)(in ? c : DD - 1) * DD + m];
                Vh[m * TC + cc] = in ? cplx<R>{(R)v.x, (R)(-v.y)} : cplx<R>{R(0), R(0)};
            }
            for (int m = tid; m < KP; m += NT) lam[m] = g.eigL[(size_t)kb * DD + m];
            cur = kb;
            __syncthreads();                    // the run's V^H and eigenvalues
        }
        // diagonal d = (ca Re f + cb Im f) + i cc Im f, f = 1 / (z - lambda_m);
        // raw output takes G itself (d = f), the panels are then made from it
        double ca = 1.0, cb = 0.0, cc = 1.0;                  // imag B: f
        if (!g.raw) {
            if (g.mode == TH_G_IMAG_A) { ca = w; cc = w; }    // w f
            if (g.mode == TH_G_REAL_A) { ca = 0.0; cb = 2.0 * w; cc = 0.0; }   // 2 w Im f
            if (g.mode == TH_G_REAL_B) cc = -1.0;             // conj(f)
        }
        for (int m = tid; m < KP; m += NT) {
            const double a = er - lam[m], n = a * a + ei * ei;
            const double fr = a / n, fi = -ei / n;
            dg[m] = {ca * fr + cb * fi, cc * fi};
        }
        __syncthreads();
        typename smfma<R>::acc accr[RPW][TB], acci[RPW][TB];
#pragma unroll
        for (int i = 0; i < RPW; ++i)
#pragma unroll
            for (int j = 0; j < TB; ++j)
                for (int q = 0; q < 4; ++q) { accr[i][j][q] = 0; acci[i][j][q] = 0; }
        // A = V[r][m] d_m from device memory (the k-point's V stays in L2 over
        // its run), loaded PF K steps ahead of its use, every load unconditional
        // (the tail re-reads the last step)
        // FP64 accumulators of two row blocks, or six FP32 blocks (d = 288, 48
        // columns), leave room for one step
        constexpr int PF = (sizeof(R) == 8 && RPW > 1) || (RPW > 1 && TB > 2) ? 1 : 4;
        static_assert(KS >= PF, "prefetch depth");
        double2 vq[PF][RPW];
        const double2 *Vl = Vk + (lane >> 4);
#pragma unroll
        for (int u = 0; u < PF; ++u)
#pragma unroll
            for (int i = 0; i < RPW; ++i) vq[u][i] = Vl[(size_t)arc[i] * DD + 4 * u];
        for (int k0 = 0; k0 < KS; k0 += PF) {
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                const int ks = k0 + u;
                if (ks < KS) {                  // the same in every thread
                    const int m = 4 * ks + (lane >> 4);
                    const cplx<double> dm = dg[m];
                    const int kn = ks + PF < KS ? ks + PF : KS - 1;
                    R ar[RPW], ai[RPW], nai[RPW];
#pragma unroll
                    for (int i = 0; i < RPW; ++i) {
                        const double2 v = vq[u][i];
                        vq[u][i] = Vl[(size_t)arc[i] * DD + 4 * kn];
                        const cplx<double> x = cmul(cplx<double>{v.x, v.y}, dm);
                        ar[i] = arv[i] ? (R)x.x : R(0);
                        ai[i] = arv[i] ? (R)x.y : R(0);
                        nai[i] = -ai[i];
                    }
#pragma unroll
                    for (int j = 0; j < TB; ++j) {
                        const cplx<R> b = Vh[m * TC + 16 * j + (lane & 15)];
#pragma unroll
                        for (int i = 0; i < RPW; ++i) {
                            accr[i][j] = smfma<R>::mac(ar[i], b.x, accr[i][j]);
                            accr[i][j] = smfma<R>::mac(nai[i], b.y, accr[i][j]);
                            acci[i][j] = smfma<R>::mac(ar[i], b.y, acci[i][j]);
                            acci[i][j] = smfma<R>::mac(ai[i], b.x, acci[i][j]);
                        }
                    }
                }
            }
        }
        if (g.panel && !g.raw) {
            P *out = reinterpret_cast<P *>(g.panel) + panel_row(g, pl) * g.nA2 * 2 * Mpad;
#pragma unroll
            for (int i = 0; i < RPW; ++i)
#pragma unroll
                for (int j = 0; j < TB; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (rv[i][q] && cv[j]) {
                            const int blk = bside ? cI[j] * nA + rI[i][q] : rI[i][q] * nA + cI[j];
                            const size_t o = (size_t)blk * 2 * Mpad + rR[i][q] + cR[j];
                            out[o] = (P)accr[i][j][q];
                            out[o + Mpad] = (P)acci[i][j][q];
                        }
            if (blockIdx.y == 0) {
                const int npad = Mpad - Mb;
                for (int e = tid; e < g.nA2 * 2 * npad; e += NT) {
                    const int part = e / npad;
                    out[(size_t)part * Mpad + Mb + (e - part * npad)] = P(0);
                }
            }
        }
        if (g.raw) {
            double2 *o = g.raw + (size_t)pl * DD * DD;
#pragma unroll
            for (int i = 0; i < RPW; ++i)
#pragma unroll
                for (int j = 0; j < TB; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (rv[i][q] && cv[j]) {
                            const int r = 16 * (wave + NW * i) + smfma<R>::row(lane >> 4, q);
                            const int c = c0 + 16 * j + (lane & 15);
                            o[(size_t)r * DD + c] = make_double2((double)accr[i][j][q], (double)acci[i][j][q]);
                        }
        }
    }
}

// Mixed precision variant of the run kernel on the fp16 matrix cores
// (v_mfma_f32_16x16x32_f16, K = 32 eigen-indices per instruction).  The
// operands are split in two fp16 parts (hi + lo, 22 significant bits):
// V^H once per run (copied from the split planes of 2^14 V kept with the
// eigenpairs, the scale keeping small entries' lo parts normal), V diag(d) per pair under a
// power-of-two scale from |d_m| <= max(|ca|, |cb|, |cc|) / Im z (the
// resolvent bound of a Hermitian H), and each real product is hi.hi +
// hi.lo + lo.hi accumulated in FP32 (Ar Br and Ai Bi in separate
// accumulators).  The diagonal d_m is computed once per pair (one eigenvalue
// per thread, double-buffered in LDS by pair parity, one barrier per pair).
typedef _Float16 ghv8 __attribute__((ext_vector_type(8)));
constexpr unsigned TH_G_OOB = 0x10000000u;     // byte offset past any panel row (dropped stores)
template <int DD, int CH>
__global__ __launch_bounds__(64 * ((DD + 15) / 16)) __attribute__((amdgpu_waves_per_eu(3))) void green_spectral_h2_kernel(green_params g)
{
    constexpr int NB = (DD + 15) / 16, NT = 64 * NB, KT = (DD + 31) / 32, KP = 32 * KT, NC = 16 * NB;
    constexpr int VST = KP + 8;                  // LDS row stride in halves (16-byte reads on distinct banks)
    __shared__ __align__(16) _Float16 Vs[2][2][NC * VST];   // [re|im][hi|lo][c][m] = conj(V[c][m])
    __shared__ float2 dv[CH][KP];                // diagonals of the workgroup's pairs
    __shared__ double pzr[CH], pzi[CH], pzw[CH]; // Re z, Im z and weight of the workgroup's pairs
    __shared__ int pkb[CH];                      // their eigen-decomposition
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const long s0 = (long)blockIdx.x * CH;
    const long s1 = s0 + CH < g.np ? s0 + CH : g.np;
    const bool imag = g.mode == TH_G_IMAG_A || g.mode == TH_G_IMAG_B;
    const bool bside = g.mode == TH_G_IMAG_B || g.mode == TH_G_REAL_B;
    const int bs = g.bs, nA = g.nA, Mpad = g.Mpad, Mb = bs * bs;
    // the pairs' scalars are read once, so that no global load sits between
    // the per-pair barriers
    if (tid < CH && s0 + tid < s1) {
        const long p = g.p0 + s0 + tid;
        double k0, k1, k2, er, ei, w;
        pair_point(g, p, k0, k1, k2, er, ei, w);
        long kb;
        if (imag) {
            kb = g.kid_im[p];
        } else {
            long ie, ik;
            real_split(g, p, ie, ik);
            kb = g.kid_re[ik];
        }
        pzr[tid] = er;
        pzi[tid] = ei;
        pzw[tid] = w;
        pkb[tid] = (int)kb;
    }
    // entry m of the diagonal d = (ca Re f + cb Im f) + i cc Im f of pair jj,
    // f = 1 / (z - lam) with lam the m-th eigenvalue of its k-point
    auto diagonal = [&](int jj, double lam) {
        const double w = pzw[jj];
        double ca = 1.0, cb = 0.0, cc = 1.0;
        if (g.mode == TH_G_IMAG_A) { ca = w; cc = w; }
        if (g.mode == TH_G_REAL_A) { ca = 0.0; cb = 2.0 * w; cc = 0.0; }
        if (g.mode == TH_G_REAL_B) cc = -1.0;
        const double ei = pzi[jj], a = pzr[jj] - lam, n = a * a + ei * ei;
        const double fr = a / n, fi = -ei / n;
        return make_float2((float)(ca * fr + cb * fi), (float)(cc * fi));
    };
    int rI[4], rR[4], cI[NB], cR[NB];
    bool rv[4], cv[NB];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 16 * wave + 4 * (lane >> 4) + q;
        rv[q] = r < DD;
        rI[q] = r / bs;
        rR[q] = (r - rI[q] * bs) * bs;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int c = 16 * j + (lane & 15);
        cv[j] = c < DD;
        cI[j] = c / bs;
        cR[j] = c - cI[j] * bs;
    }
    const int arow = 16 * wave + (lane & 15);   // row of V in this lane's A operand
    const int kofs = 8 * (lane >> 4);            // first of the lane's 8 eigen-indices per K step
    // byte offsets of the lane's outputs in a pair's panel row, a row part plus
    // a column part; an entry outside the matrix gets an offset far past the
    // row (TH_G_OOB >= the row's bytes, checked by the launcher), which the
    // buffer store drops
    unsigned rof[4], cof[NB];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        rof[q] = rv[q] ? 4u * (unsigned)((bside ? rI[q] : rI[q] * nA) * 2 * Mpad + rR[q]) : TH_G_OOB;
#pragma unroll
    for (int j = 0; j < NB; ++j)
        cof[j] = cv[j] ? 4u * (unsigned)((bside ? cI[j] * nA : cI[j]) * 2 * Mpad + cR[j]) : TH_G_OOB;
    const int rowbytes = 4 * g.nA2 * 2 * Mpad, imoff = 4 * Mpad;
    // split panels: exponent of the panel (the real-axis A side from the
    // current node weights)
    const int eP = (g.presplit && g.pareal > 0.0) ? th_real_exp(g.p2, g.pnx, g.pareal) : g.pexp;
    long cur = -1;
    float avr[KT][8], avi[KT][8];
    float bmax = 0.0f;                          // largest |entry| written (FP32 panels)
    __syncthreads();                            // the pairs' scalars
    for (long pl = s0; pl < s1; ++pl) {
        const int jj = (int)(pl - s0);
        const long kb = pkb[jj];
        const double ei = pzi[jj], w = pzw[jj];
        // diagonal coefficients: d = (ca Re f + cb Im f) + i cc Im f
        double ca = 1.0, cb = 0.0, cc = 1.0;
        if (g.mode == TH_G_IMAG_A) { ca = w; cc = w; }
        if (g.mode == TH_G_REAL_A) { ca = 0.0; cb = 2.0 * w; cc = 0.0; }
        if (g.mode == TH_G_REAL_B) cc = -1.0;
        if (kb != cur) {                        // same kb in every thread
            __syncthreads();                    // the previous run's reads of Vs and dv are done
            // V^H and the A rows from the k-point's split planes of 2^14 V (rows
            // and eigen-indices padded with zeros): unconditional 16-byte loads,
            // all issued before the first wait
            constexpr int PL = NC * KP / 8;      // 8-half vectors per plane
            const ghv8 *Sk = reinterpret_cast<const ghv8 *>(g.eigS) + (size_t)kb * 4 * PL;
            for (int e = tid; e < PL; e += NT) {
                const int c = e / (KP / 8), m0 = 8 * (e - c * (KP / 8));
                const ghv8 rh = Sk[e], rl = Sk[PL + e], ih = Sk[2 * PL + e], il = Sk[3 * PL + e];
                *reinterpret_cast<ghv8 *>(&Vs[0][0][c * VST + m0]) = rh;
                *reinterpret_cast<ghv8 *>(&Vs[0][1][c * VST + m0]) = rl;
                *reinterpret_cast<ghv8 *>(&Vs[1][0][c * VST + m0]) = -ih;   // conj
                *reinterpret_cast<ghv8 *>(&Vs[1][1][c * VST + m0]) = -il;
            }
            const float us = ldexpf(1.0f, -14);
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int o = (arow * KP + 32 * t + kofs) / 8;
                const ghv8 rh = Sk[o], rl = Sk[PL + o], ih = Sk[2 * PL + o], il = Sk[3 * PL + o];
#pragma unroll
                for (int i = 0; i < 8; ++i) {   // V to 22 bits
                    avr[t][i] = ((float)rh[i] + (float)rl[i]) * us;
                    avi[t][i] = ((float)ih[i] + (float)il[i]) * us;
                }
            }
            // the diagonals of every later pair of the workgroup on this k-point,
            // so that the run needs no barrier per pair (eigenvalue loads ahead of
            // the branch)
            const double *Lk = g.eigL + (size_t)kb * DD;
            for (int e = tid; e < CH * KP; e += NT) {
                const int j = e / KP, m = e - (e / KP) * KP;
                const double lam = Lk[m < DD ? m : DD - 1];
                if (j >= jj && s0 + j < s1 && pkb[j] == kb)
                    dv[j][m] = m < DD ? diagonal(j, lam) : make_float2(0.0f, 0.0f);
            }
            cur = kb;
            __syncthreads();                    // the run's V and diagonals in LDS
        }
        // power-of-two scale of V diag(d): |d_m| <= max(|ca|, |cb|, |cc|) / Im z
        const double bound = fmax(fmax(fabs(ca), fabs(cb)), fabs(cc)) / fabs(ei);
        // bound = m 2^ex with m in [0.5, 1) (exponent bits; an infinite or zero
        // bound clamps to the smallest / largest scale)
        const int ex = (int)((__double_as_longlong(bound) >> 52) & 0x7ff) - 1022;
        const int es = min(100, max(-100, 14 - ex));
        const float sc = ldexpf(1.0f, es);
        ghv8 arh[KT], arl[KT], aih[KT], ail[KT];
#pragma unroll
        for (int t = 0; t < KT; ++t)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float2 d = dv[jj][32 * t + kofs + i];
                // x sc (a power of two) and x sc - hi are exact in FP32: one
                // fused multiply-add rounded once to fp16 gives each part
                const float xr = avr[t][i] * d.x - avi[t][i] * d.y;
                const float xi = avr[t][i] * d.y + avi[t][i] * d.x;
                const _Float16 hr = (_Float16)fmaf(xr, sc, 0.0f), hi_ = (_Float16)fmaf(xi, sc, 0.0f);
                arh[t][i] = hr;
                arl[t][i] = (_Float16)fmaf(xr, sc, -(float)hr);
                aih[t][i] = hi_;
                ail[t][i] = (_Float16)fmaf(xi, sc, -(float)hi_);
            }
        sfv4 acp[NB], acn[NB], aci[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j)
            for (int q = 0; q < 4; ++q) { acp[j][q] = 0.0f; acn[j][q] = 0.0f; aci[j][q] = 0.0f; }
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int bo = (16 * j + (lane & 15)) * VST + 32 * t + kofs;
                const ghv8 brh = *reinterpret_cast<const ghv8 *>(&Vs[0][0][bo]);
                const ghv8 brl = *reinterpret_cast<const ghv8 *>(&Vs[0][1][bo]);
                const ghv8 bih = *reinterpret_cast<const ghv8 *>(&Vs[1][0][bo]);
                const ghv8 bil = *reinterpret_cast<const ghv8 *>(&Vs[1][1][bo]);
                acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], brh, acp[j], 0, 0, 0);
                acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], bih, acn[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], bih, aci[j], 0, 0, 0);
                acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], brl, acp[j], 0, 0, 0);
                acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], bil, acn[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], bil, aci[j], 0, 0, 0);
                acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[t], brh, acp[j], 0, 0, 0);
                acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[t], bih, acn[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[t], bih, aci[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], brh, aci[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], brl, aci[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[t], brh, aci[j], 0, 0, 0);
            }
        // accumulators = 2^(es + 14) G (A carries 2^es, V^H from the planes 2^14)
        const int eo = es + 14;
        const float back = ldexpf(1.0f, -eo);
        const int npad = Mpad - Mb;
        void *row = reinterpret_cast<unsigned *>(g.panel) + panel_row(g, pl) * g.nA2 * 2 * Mpad;
        const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(row, (short)0, rowbytes, 0x00020000);
        auto put = [&](int j, int q, unsigned wr, unsigned wi) {
            const unsigned o = rof[q] + cof[j];
            __builtin_amdgcn_raw_buffer_store_b32(wr, ro, o, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(wi, ro, o, imoff, 0);
        };
        auto pack = [](_Float16 h, _Float16 l) {   // hi in the low half
            typedef _Float16 h2v __attribute__((ext_vector_type(2)));
            const h2v v = {h, l};
            return __builtin_bit_cast(unsigned, v);
        };
        if (g.presplit && eP - eo >= -126 && eP - eo <= 127) {
            // x 2^eP = hi + lo as the contraction would split the FP32 value
            // x = acc 2^-es; y 2^(eP-es) (y the accumulator, 2^(eP-es) a normal
            // float) and y 2^(eP-es) - hi are exact in FP32, so one fused
            // multiply-add rounded once to fp16 gives each part; hi and lo share
            // one word
            const float fs = ldexpf(1.0f, eP - eo);
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float yr = acp[j][q] - acn[j][q], yi = aci[j][q];
                    const _Float16 hr = (_Float16)fmaf(yr, fs, 0.0f), hi_ = (_Float16)fmaf(yi, fs, 0.0f);
                    const _Float16 lr = (_Float16)fmaf(yr, fs, -(float)hr), li = (_Float16)fmaf(yi, fs, -(float)hi_);
                    put(j, q, pack(hr, lr), pack(hi_, li));
                }
        } else if (g.presplit) {
            // the same split with the scaling by ldexpf (2^(eP-es) outside the
            // normal floats)
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float xr = ldexpf(acp[j][q] - acn[j][q], eP - eo), xi = ldexpf(aci[j][q], eP - eo);
                    const _Float16 hr = (_Float16)xr, hi_ = (_Float16)xi;
                    const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi_);
                    put(j, q, pack(hr, lr), pack(hi_, li));
                }
        } else {
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float vr = (acp[j][q] - acn[j][q]) * back, vi = aci[j][q] * back;
                    bmax = fmaxf(bmax, fmaxf(fabsf(vr), fabsf(vi)));   // padding entries hold zeros
                    put(j, q, __builtin_bit_cast(unsigned, vr), __builtin_bit_cast(unsigned, vi));
                }
        }
        for (int e = tid; e < g.nA2 * 2 * npad; e += NT) {
            const int part = e / npad;
            __builtin_amdgcn_raw_buffer_store_b32(0u, ro, 4u * (unsigned)(part * Mpad + Mb + (e - part * npad)), 0, 0);
        }
    }
    if (g.blkmax && !g.presplit) {              // the workgroup's largest |entry|
        __shared__ float wmax[NB];
        for (int o = 32; o > 0; o >>= 1) bmax = fmaxf(bmax, __shfl_xor(bmax, o));
        if (lane == 0) wmax[wave] = bmax;
        __syncthreads();
        if (tid == 0) {
            for (int w2 = 1; w2 < NB; ++w2) bmax = fmaxf(bmax, wmax[w2]);
            g.blkmax[blockIdx.x] = bmax;
        }
    }
}


// Column-tiled spectral kernel on the fp16 matrix cores for d = 72 and 144
// (mixed precision): G = V (diag(d) V^H) for a run of up to CH pairs on one
// k-point, one workgroup per (run, tile of TC columns), one wave per 16-row
// block.  The diagonal rides on the B side, so only the B tile depends on the
// pair: A = 2^SA V (the wave's 16 rows, every eigen-index, two-term fp16
// split, the scale keeping the lo parts of small entries normal) stays in
// registers for the run, and so do each thread's groups of 8 eigen-indices of
// conj(V) for the tile, in FP32.  Per pair a thread multiplies its groups by
// 2^es d_m (the pair's diagonal with the power-of-two scale of the resolvent
// bound folded in), splits them and writes the B tile into LDS (two tiles by
// pair parity, one barrier per pair).  Products and output as in the run
// kernel: hi.hi + hi.lo + lo.hi per real product, Ar Br and Ai Bi in separate
// accumulators, panels as split words or FP32.
template <int DD, int TC, int CH>
__global__ __launch_bounds__(64 * ((DD + 15) / 16)) __attribute__((amdgpu_waves_per_eu(3))) void
green_spectral_tiled_h2_kernel(green_params g)
{
    constexpr int NB = (DD + 15) / 16, NT = 64 * NB, KT = (DD + 31) / 32, KP = 32 * KT, TB = TC / 16;
    constexpr int VST = KP + 8;                  // LDS row stride in halves (16-byte reads on distinct banks)
    constexpr int GR = KP / 8, NG = TC * GR, GPT = (NG + NT - 1) / NT;   // 8-index groups of the tile
    constexpr int SA = 14;                       // |2^SA V| <= 2^14 < the fp16 maximum
    static_assert(TC % 16 == 0, "column tile of whole 16-wide blocks");
    __shared__ __align__(16) _Float16 Bs[2][4][TC * VST];   // [pair parity][re hi|re lo|im hi|im lo][c][m]
    __shared__ __align__(16) float2 dv[CH][KP];              // 2^es d of the workgroup's pairs
    __shared__ double pzr[CH], pzi[CH], pzw[CH];
    __shared__ int pkb[CH], pes[CH];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int c0 = (int)blockIdx.y * TC;
    const long s0 = (long)blockIdx.x * CH;
    const long s1 = s0 + CH < g.np ? s0 + CH : g.np;
    const bool imag = g.mode == TH_G_IMAG_A || g.mode == TH_G_IMAG_B;
    const bool bside = g.mode == TH_G_IMAG_B || g.mode == TH_G_REAL_B;
    const int bs = g.bs, nA = g.nA, Mpad = g.Mpad, Mb = bs * bs;
    // coefficients of d = (ca Re f + cb Im f) + i cc Im f, f = 1 / (z - lam)
    auto coeffs = [&](double w, double &ca, double &cb, double &cc) {
        ca = 1.0; cb = 0.0; cc = 1.0;
        if (g.mode == TH_G_IMAG_A) { ca = w; cc = w; }
        if (g.mode == TH_G_REAL_A) { ca = 0.0; cb = 2.0 * w; cc = 0.0; }
        if (g.mode == TH_G_REAL_B) cc = -1.0;
    };
    if (tid < CH && s0 + tid < s1) {
        const long p = g.p0 + s0 + tid;
        double k0, k1, k2, er, ei, w;
        pair_point(g, p, k0, k1, k2, er, ei, w);
        long kb;
        if (imag) {
            kb = g.kid_im[p];
        } else {
            long ie, ik;
            real_split(g, p, ie, ik);
            kb = g.kid_re[ik];
        }
        pzr[tid] = er;
        pzi[tid] = ei;
        pzw[tid] = w;
        pkb[tid] = (int)kb;
        // power-of-two scale of d: |d_m| <= max(|ca|, |cb|, |cc|) / Im z = m 2^ex,
        // m in [0.5, 1), so |2^es d_m| < 2^14 (an infinite or zero bound clamps)
        double ca, cb, cc;
        coeffs(w, ca, cb, cc);
        const double bound = fmax(fmax(fabs(ca), fabs(cb)), fabs(cc)) / fabs(ei);
        const int ex = (int)((__double_as_longlong(bound) >> 52) & 0x7ff) - 1022;
        pes[tid] = min(100, max(-100, 14 - ex));
    }
    int rI[4], rR[4], cI[TB], cR[TB];
    bool rv[4], cv[TB];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 16 * wave + 4 * (lane >> 4) + q;
        rv[q] = r < DD;
        rI[q] = r / bs;
        rR[q] = (r - rI[q] * bs) * bs;
    }
#pragma unroll
    for (int j = 0; j < TB; ++j) {
        const int c = c0 + 16 * j + (lane & 15);
        cv[j] = c < DD;
        cI[j] = c / bs;
        cR[j] = c - cI[j] * bs;
    }
    // valid 16-column blocks of this tile (the same in every thread)
    const int nbv = min(TB, (DD - c0 + 15) / 16);
    const int arow = 16 * wave + (lane & 15);   // row of V in this lane's A operand
    const int kofs = 8 * (lane >> 4);            // first of the lane's 8 eigen-indices per K step
    unsigned rof[4], cof[TB];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        rof[q] = rv[q] ? 4u * (unsigned)((bside ? rI[q] : rI[q] * nA) * 2 * Mpad + rR[q]) : TH_G_OOB;
#pragma unroll
    for (int j = 0; j < TB; ++j)
        cof[j] = cv[j] ? 4u * (unsigned)((bside ? cI[j] * nA : cI[j]) * 2 * Mpad + cR[j]) : TH_G_OOB;
    const int rowbytes = 4 * g.nA2 * 2 * Mpad, imoff = 4 * Mpad;
    const int eP = (g.presplit && g.pareal > 0.0) ? th_real_exp(g.p2, g.pnx, g.pareal) : g.pexp;
    ghv8 arh[KT], arl[KT], aih[KT], ail[KT];
    float bsr[GPT][8], bsi[GPT][8];
    // B values of this thread's groups for the next pair, built while the
    // current pair's products run
    ghv8 nbh[GPT], nbl[GPT], nih[GPT], nil[GPT];
    // element i of group s for pair j: 2^es d_m conj(V[c][m]) = hi + lo (the lo
    // part from one fused multiply-add of the exact FP32 remainder); a group
    // past the tile reads row j of dv in range and is never stored
    auto prep = [&](int j, int s, int i) {
        const int gi = tid + NT * s, m0 = 8 * (gi - (gi / GR) * GR);
        const float2 d = dv[j][m0 + i];
        const float xr = fmaf(d.x, bsr[s][i], -d.y * bsi[s][i]);
        const float xi = fmaf(d.x, bsi[s][i], d.y * bsr[s][i]);
        const _Float16 hr = (_Float16)xr, hi_ = (_Float16)xi;
        nbh[s][i] = hr;
        nbl[s][i] = (_Float16)fmaf(xr, 1.0f, -(float)hr);
        nih[s][i] = hi_;
        nil[s][i] = (_Float16)fmaf(xi, 1.0f, -(float)hi_);
    };
    auto store = [&](int par) {
#pragma unroll
        for (int s = 0; s < GPT; ++s) {
            const int gi = tid + NT * s;
            if (gi < NG) {
                const int cl = gi / GR, o = cl * VST + 8 * (gi - cl * GR);
                *reinterpret_cast<ghv8 *>(&Bs[par][0][o]) = nbh[s];
                *reinterpret_cast<ghv8 *>(&Bs[par][1][o]) = nbl[s];
                *reinterpret_cast<ghv8 *>(&Bs[par][2][o]) = nih[s];
                *reinterpret_cast<ghv8 *>(&Bs[par][3][o]) = nil[s];
            }
        }
    };
    constexpr int PE = (8 + KT - 1) / KT;       // next-pair elements per K step
    long cur = -1;
    __syncthreads();                            // the pairs' scalars
    for (long pl = s0; pl < s1; ++pl) {
        const int jj = (int)(pl - s0);
        const long kb = pkb[jj];
        if (kb != cur) {                        // same kb in every thread; the last
                                                // pair ended with a barrier
            // A and the B sources from the k-point's split planes of 2^SA V (rows
            // and eigen-indices padded with zeros): sixteen bytes per plane and
            // load, straight into the operand registers
            constexpr int RP = 16 * NB, PL = RP * KP / 8;       // 8-half vectors per plane
            const ghv8 *Sk = reinterpret_cast<const ghv8 *>(g.eigS) + (size_t)kb * 4 * PL;
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int o = (arow * KP + 32 * t + kofs) / 8;
                arh[t] = Sk[o];
                arl[t] = Sk[PL + o];
                aih[t] = Sk[2 * PL + o];
                ail[t] = Sk[3 * PL + o];
            }
            const float us = ldexpf(1.0f, -SA);
#pragma unroll
            for (int s = 0; s < GPT; ++s) {
                const int gi = tid + NT * s, cl = gi / GR, m0 = 8 * (gi - cl * GR);
                const bool in = gi < NG;                        // rows c0 + cl < RP
                const int o = ((in ? c0 + cl : 0) * KP + m0) / 8;
                const ghv8 rh = Sk[o], rl = Sk[PL + o], ih = Sk[2 * PL + o], il = Sk[3 * PL + o];
#pragma unroll
                for (int i = 0; i < 8; ++i) {   // conj(V[c][m]) to 22 bits
                    bsr[s][i] = in ? ((float)rh[i] + (float)rl[i]) * us : 0.0f;
                    bsi[s][i] = in ? -((float)ih[i] + (float)il[i]) * us : 0.0f;
                }
            }
            // 2^es d for every later pair of the workgroup on this k-point
            const double *Lk = g.eigL + (size_t)kb * DD;
            for (int e = tid; e < CH * KP; e += NT) {
                const int j = e / KP, m = e - (e / KP) * KP;
                const double lam = Lk[m < DD ? m : DD - 1];
                if (j >= jj && s0 + j < s1 && pkb[j] == kb) {
                    float2 v = make_float2(0.0f, 0.0f);
                    if (m < DD) {
                        double ca, cb, cc;
                        coeffs(pzw[j], ca, cb, cc);
                        const double ei = pzi[j], a = pzr[j] - lam, n = a * a + ei * ei;
                        const double fr = a / n, fi = -ei / n;
                        v = make_float2((float)ldexp(ca * fr + cb * fi, pes[j]), (float)ldexp(cc * fi, pes[j]));
                    }
                    dv[j][m] = v;
                }
            }
            cur = kb;
            __syncthreads();                    // the run's diagonals
#pragma unroll
            for (int s = 0; s < GPT; ++s)
#pragma unroll
                for (int i = 0; i < 8; ++i) prep(jj, s, i);
            store(jj & 1);                      // last read two pairs ago
            __syncthreads();                    // the run's first B tile
        }
        // B of this pair is in tile par; B of the next pair on this k-point goes
        // to the other tile (last read by the previous pair, whose barrier every
        // thread has passed)
        const int par = jj & 1;
        const bool nxt = pl + 1 < s1 && pkb[jj + 1] == kb;
        sfv4 acp[TB], acn[TB], aci[TB];
#pragma unroll
        for (int j = 0; j < TB; ++j)
            for (int q = 0; q < 4; ++q) { acp[j][q] = 0.0f; acn[j][q] = 0.0f; aci[j][q] = 0.0f; }
#pragma unroll
        for (int j = 0; j < TB; ++j) {
            if (j >= nbv) break;
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int bo = (16 * j + (lane & 15)) * VST + 32 * t + kofs;
                const ghv8 brh = *reinterpret_cast<const ghv8 *>(&Bs[par][0][bo]);
                const ghv8 brl = *reinterpret_cast<const ghv8 *>(&Bs[par][1][bo]);
                const ghv8 bih = *reinterpret_cast<const ghv8 *>(&Bs[par][2][bo]);
                const ghv8 bil = *reinterpret_cast<const ghv8 *>(&Bs[par][3][bo]);
                acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], brh, acp[j], 0, 0, 0);
                acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], bih, acn[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], bih, aci[j], 0, 0, 0);
                acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], brl, acp[j], 0, 0, 0);
                acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], bil, acn[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh[t], bil, aci[j], 0, 0, 0);
                acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[t], brh, acp[j], 0, 0, 0);
                acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[t], bih, acn[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl[t], bih, aci[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], brh, aci[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih[t], brl, aci[j], 0, 0, 0);
                aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail[t], brh, aci[j], 0, 0, 0);
                if (j == 0 && nxt) {            // the next pair's B, a few elements per step
#pragma unroll
                    for (int s = 0; s < GPT; ++s)
#pragma unroll
                        for (int e = 0; e < PE; ++e)
                            if (PE * t + e < 8) prep(jj + 1, s, PE * t + e);
                }
            }
        }
        if (nxt) store(par ^ 1);
        // accumulators = 2^(es + SA) G
        const int es = pes[jj] + SA;
        void *row = reinterpret_cast<unsigned *>(g.panel) + panel_row(g, pl) * g.nA2 * 2 * Mpad;
        const __amdgpu_buffer_rsrc_t ro = __builtin_amdgcn_make_buffer_rsrc(row, (short)0, rowbytes, 0x00020000);
        auto put = [&](int j, int q, unsigned wr, unsigned wi) {
            const unsigned o = rof[q] + cof[j];
            __builtin_amdgcn_raw_buffer_store_b32(wr, ro, o, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(wi, ro, o, imoff, 0);
        };
        auto pack = [](_Float16 h, _Float16 l) {   // hi in the low half
            typedef _Float16 h2v __attribute__((ext_vector_type(2)));
            const h2v v = {h, l};
            return __builtin_bit_cast(unsigned, v);
        };
        if (g.presplit && eP - es >= -126 && eP - es <= 127) {
            // y 2^(eP-es) and y 2^(eP-es) - hi are exact in FP32: one fused
            // multiply-add rounded once to fp16 gives each part of x 2^eP
            const float fs = ldexpf(1.0f, eP - es);
#pragma unroll
            for (int j = 0; j < TB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float yr = acp[j][q] - acn[j][q], yi = aci[j][q];
                    const _Float16 hr = (_Float16)fmaf(yr, fs, 0.0f), hi_ = (_Float16)fmaf(yi, fs, 0.0f);
                    const _Float16 lr = (_Float16)fmaf(yr, fs, -(float)hr), li = (_Float16)fmaf(yi, fs, -(float)hi_);
                    put(j, q, pack(hr, lr), pack(hi_, li));
                }
        } else if (g.presplit) {
#pragma unroll
            for (int j = 0; j < TB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float xr = ldexpf(acp[j][q] - acn[j][q], eP - es), xi = ldexpf(aci[j][q], eP - es);
                    const _Float16 hr = (_Float16)xr, hi_ = (_Float16)xi;
                    const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi_);
                    put(j, q, pack(hr, lr), pack(hi_, li));
                }
        } else {
            const float back = ldexpf(1.0f, -es);
#pragma unroll
            for (int j = 0; j < TB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    put(j, q, __builtin_bit_cast(unsigned, (acp[j][q] - acn[j][q]) * back),
                        __builtin_bit_cast(unsigned, aci[j][q] * back));
        }
        if (blockIdx.y == 0) {
            const int npad = Mpad - Mb;
            for (int e = tid; e < g.nA2 * 2 * npad; e += NT) {
                const int part = e / npad;
                __builtin_amdgcn_raw_buffer_store_b32(0u, ro, 4u * (unsigned)(part * Mpad + Mb + (e - part * npad)), 0,
                                                      0);
            }
        }
        __syncthreads();                        // the next pair's B tile; this pair's reads done
    }
}


template <typename P, int NT>
__global__ __launch_bounds__(NT) void green_kernel(green_params g)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int d = g.d, tid = threadIdx.x;
    const long pl = blockIdx.x;              // pair within the batch
    const long p = g.p0 + pl;                // global pair index
    double k0, k1, k2, er, ei, w;
    pair_point(g, p, k0, k1, k2, er, ei, w);
    // ---- LDS layout ----
    cplx<double> *M = reinterpret_cast<cplx<double> *>(smem);          // d*d, FP64 (z - H)
    size_t off = sizeof(cplx<double>) * d * d;
    cplx<double> *ph = reinterpret_cast<cplx<double> *>(smem + off);    // nhop phases
    off += sizeof(cplx<double>) * g.nhop;
    cplx<double> *colk = reinterpret_cast<cplx<double> *>(smem + off);
    off += sizeof(cplx<double>) * d;
    cplx<double> *prow = reinterpret_cast<cplx<double> *>(smem + off);
    off += sizeof(cplx<double>) * d;
    int *perm = reinterpret_cast<int *>(smem + off);
    off += sizeof(int) * d;
    int *idx = reinterpret_cast<int *>(smem + off);
    off += sizeof(int) * d;
    int *s_piv = reinterpret_cast<int *>(smem + off);
    off += 16;
    cplx<double> *s_inv = reinterpret_cast<cplx<double> *>(smem + off);
    off += 16;

    // ---- z - H(k) ----
    for (int e = tid; e < d * d; e += NT) {
        int r = e / d, c = e - r * d;
        double2 hv = g.h0[r + (size_t)c * d];
        M[e] = {(r == c ? er : 0.0) - hv.x, (r == c ? ei : 0.0) - hv.y};
    }
    for (int n = tid; n < g.nhop; n += NT) {
        double a = k0 * g.R[3 * n] + k1 * g.R[3 * n + 1] + k2 * g.R[3 * n + 2];
        double s, c;
        sincos(a, &s, &c);
        ph[n] = {c, s};
    }
    __syncthreads();
    {
        // element (nu, mu) of every hop block is always handled by thread
        // e % NT, so hops that land on the same entry never race
        const int nO = g.nOrb, nO2 = nO * nO;
        for (int e = tid; e < nO2; e += NT) {
            const int nu = e % nO, mu = e / nO;   // t(nu, mu), rows = orbitals of jb
            for (int n = 0; n < g.nhop; ++n) {
                double2 tv = g.t[(size_t)n * nO2 + nu + nO * mu];
                cplx<double> v = cmul(cplx<double>{tv.x, tv.y}, ph[n]);
                int j = g.jb[n], i = g.ib[n];
                for (int s = 0; s < 2; ++s) {
                    int row = (2 * j + s) * nO + nu, col = (2 * i + s) * nO + mu;
                    M[row * d + col] = csub(M[row * d + col], v);
                }
            }
        }
    }
    __syncthreads();
    // ---- inverse ----
    gj_invert<double, NT>(M, d, colk, prow, perm, idx, s_piv, s_inv, g.status);
    if (g.raw) {
        double2 *o = g.raw + (size_t)pl * d * d;
        for (int e = tid; e < d * d; e += NT) {
            const int r = e / d, c = e - (e / d) * d;
            const cplx<double> x = M[r * d + idx[c]];
            o[e] = make_double2(x.x, x.y);
        }
    }
    write_panel<P, NT>(g, pl, M, idx, d, w);
}

inline hipStream_t gstream(th_handle *h, const th_green_args &a) { return a.stream ? a.stream : h->stream; }

template <typename P, int NT>
int launch_green_t(th_handle *h, const th_green_args &a, const green_params &g);

template <typename P, int DD, int NT>
int launch_green_reg(th_handle *h, const th_green_args &a, const green_params &g)
{
    const bool full = (DD * DD * 16) <= 49152;
    size_t lds = sizeof(double2) * ((full ? (size_t)DD * DD : (size_t)2 * h->bs * h->bs) + h->nhop);
    hipLaunchKernelGGL((green_reg_kernel<P, DD, NT>), dim3((unsigned)a.np), dim3(NT), lds, gstream(h, a), g);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

template <typename P, int DD, int TY, int TX>
int launch_green_2d(th_handle *h, const th_green_args &a, const green_params &g)
{
    const bool full = full_stage(DD, TY * TX);
    size_t lds = sizeof(double2) * ((full ? (size_t)DD * DD : (size_t)2 * h->bs * h->bs) + h->nhop);
    hipLaunchKernelGGL((green_2d_kernel<P, DD, TY, TX>), dim3((unsigned)a.np), dim3(TY * TX), lds, gstream(h, a), g);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

// G of a batch by 2 x 2 block elimination (d = 288), in chunks of pairs so
// that the work space stays bounded
template <typename P>
int launch_green_block(th_handle *h, const th_green_args &a, const green_params &g0)
{
    const int d = h->d, b = d / 2;
    const size_t dd = (size_t)d * d, bb = (size_t)b * b;
    constexpr long CHUNK = 64;
    const size_t need = (size_t)CHUNK * (2 * dd + 4 * bb) * sizeof(double2);
    if (need > h->blk_bytes) {
        hipFree(h->d_blk);
        h->d_blk = nullptr;
        h->blk_bytes = 0;
        TH_HIP(hipMalloc(&h->d_blk, need));
        h->blk_bytes = need;
    }
    double2 *M = reinterpret_cast<double2 *>(h->d_blk);
    double2 *G = M + CHUNK * dd;
    double2 *Ai = G + CHUNK * dd, *X = Ai + CHUNK * bb, *Y = X + CHUNK * bb, *S = Y + CHUNK * bb;
    const int nt = (b + 31) / 32;
    auto gemm = [&](long np, double alpha, const double2 *A_, long sa, int lda, const double2 *B_, long sb, int ldb,
                    double beta, const double2 *C_, long sc, int ldc, double2 *O_, long so, int ldo) {
        hipLaunchKernelGGL(zgemm_batched_kernel, dim3(nt * nt, (unsigned)np), dim3(256), 0, gstream(h, a), b, b, b, alpha,
                           A_, sa, lda, B_, sb, ldb, beta, C_, sc, ldc, O_, so, ldo, nt);
    };
    for (long c0 = 0; c0 < a.np; c0 += CHUNK) {
        const long np = std::min(CHUNK, a.np - c0);
        green_params g = g0;
        g.p0 = a.p0 + c0;
        g.slot0 = c0;
        g.raw = a.raw ? a.raw + (size_t)c0 * dd : nullptr;
        hipLaunchKernelGGL(assemble_m_kernel, dim3((unsigned)np), dim3(256), sizeof(double2) * std::max(1, h->nhop),
                           gstream(h, a), g, M);
        // Ai = A^-1
        hipLaunchKernelGGL((inv_block_kernel<144, 16, 32>), dim3((unsigned)np), dim3(512), 0, gstream(h, a),
                           (const double2 *)M, (long)dd, d, Ai, (long)bb, b, g.status);
        // X = Ai B, Y = C Ai
        gemm(np, 1.0, Ai, bb, b, M + b, dd, d, 0.0, nullptr, 0, 0, X, bb, b);
        gemm(np, 1.0, M + (size_t)b * d, dd, d, Ai, bb, b, 0.0, nullptr, 0, 0, Y, bb, b);
        // S = E - C X
        gemm(np, -1.0, M + (size_t)b * d, dd, d, X, bb, b, 1.0, M + (size_t)b * d + b, dd, d, S, bb, b);
        // G_BB = S^-1
        hipLaunchKernelGGL((inv_block_kernel<144, 16, 32>), dim3((unsigned)np), dim3(512), 0, gstream(h, a),
                           (const double2 *)S, (long)bb, b, G + (size_t)b * d + b, (long)dd, d, g.status);
        // G_AB = -X G_BB, G_BA = -G_BB Y, G_AA = Ai - X G_BA
        gemm(np, -1.0, X, bb, b, G + (size_t)b * d + b, dd, d, 0.0, nullptr, 0, 0, G + b, dd, d);
        gemm(np, -1.0, G + (size_t)b * d + b, dd, d, Y, bb, b, 0.0, nullptr, 0, 0, G + (size_t)b * d, dd, d);
        gemm(np, -1.0, X, bb, b, G + (size_t)b * d, dd, d, 1.0, Ai, bb, b, G, dd, d);
        hipLaunchKernelGGL((panel_from_global_kernel<P, 256>), dim3((unsigned)np), dim3(256), 0, gstream(h, a), g,
                           (const double2 *)G);
        TH_HIP(hipGetLastError());
    }
    return TH_OK;
}

template <typename P, int DD>
int launch_green_spectral(th_handle *h, const th_green_args &a, const green_params &g)
{
    if (a.presplit && !(sizeof(P) == sizeof(float) && !a.raw && h->spectral_chunk > 0 && h->green_f16 && g.eigS)) {
        th_set_error("split half-precision panels come from the half-precision spectral kernel only");
        return TH_E_STATE;
    }
    if (!a.raw && h->spectral_chunk > 0) {   // runs of pairs per k-point; G is not kept
        constexpr int CH = TH_RUN, NT = 64 * ((DD + 15) / 16);
        const unsigned nblk = (unsigned)((a.np + CH - 1) / CH);
        if (sizeof(P) == sizeof(float) && h->green_f16 && g.eigS) {
            if ((size_t)4 * g.nA2 * 2 * g.Mpad >= TH_G_OOB) {
                th_set_error("half-precision spectral kernel: panel row too large");
                return TH_E_ARG;
            }
            hipLaunchKernelGGL((green_spectral_h2_kernel<DD, CH>), dim3(nblk), dim3(NT), 0, gstream(h, a), g);
            if (g.blkmax && !a.presplit) h->last_blkmax = nblk;
        } else {
            hipLaunchKernelGGL((green_spectral_chunk_kernel<P, DD, CH>), dim3(nblk), dim3(NT), 0, gstream(h, a), g);
        }
    } else {
        hipLaunchKernelGGL((green_spectral_kernel<P, DD>), dim3((unsigned)a.np), dim3(256), 0, gstream(h, a), g);
    }
    TH_HIP(hipGetLastError());
    return TH_OK;
}

template <typename P, int DD, int TC, int RPW>
int launch_green_spectral_tiled(th_handle *h, const th_green_args &a, const green_params &g)
{
    if (a.presplit) {
        th_set_error("split half-precision panels come from the half-precision spectral kernel only");
        return TH_E_STATE;
    }
    constexpr int CH = TH_RUN, NT = 64 * (((DD + 15) / 16 + RPW - 1) / RPW);
    const dim3 grid((unsigned)((a.np + CH - 1) / CH), (unsigned)((DD + TC - 1) / TC));
    hipLaunchKernelGGL((green_spectral_tiled_kernel<P, DD, TC, RPW, CH>), grid, dim3(NT), 0, gstream(h, a), g);
    TH_HIP(hipGetLastError());
    if (g.raw && g.panel) {   // the panels of the pair's mode from the raw G just written
        green_params gp = g;
        gp.raw = nullptr;
        hipLaunchKernelGGL((panel_from_global_kernel<P, 256>), dim3((unsigned)a.np), dim3(256), 0, gstream(h, a), gp,
                           (const double2 *)g.raw);
        TH_HIP(hipGetLastError());
    }
    return TH_OK;
}

// the fp16 column-tiled kernel (mixed precision, panels as the run kernel
// writes them; raw G and the per-pair variant stay with the FP32 kernel)
template <int DD, int TC>
int launch_green_spectral_tiled_h2(th_handle *h, const th_green_args &a, const green_params &g)
{
    if ((size_t)4 * g.nA2 * 2 * g.Mpad >= TH_G_OOB) {
        th_set_error("half-precision spectral kernel: panel row too large");
        return TH_E_ARG;
    }
    constexpr int CH = TH_RUN, NT = 64 * ((DD + 15) / 16);
    const dim3 grid((unsigned)((a.np + CH - 1) / CH), (unsigned)((DD + TC - 1) / TC));
    hipLaunchKernelGGL((green_spectral_tiled_h2_kernel<DD, TC, CH>), grid, dim3(NT), 0, gstream(h, a), g);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

template <typename P>
int launch_green_p(th_handle *h, const th_green_args &a, const green_params &g)
{
    if (a.presplit && (h->kernel_variant != 0 || !g.eigV ||
                       (h->d != 18 && h->d != 36 && h->d != 54 && h->d != 72 && h->d != 144))) {
        th_set_error("split half-precision panels come from the half-precision spectral kernel only");
        return TH_E_STATE;
    }
    if (h->kernel_variant == 1 && !a.raw) {   // one-dimensional ownership (kept for comparison)
        switch (h->d) {
        case 18: return launch_green_reg<P, 18, 64>(h, a, g);
        case 54: return launch_green_reg<P, 54, 256>(h, a, g);
        default: break;
        }
    }
    if (h->kernel_variant == 2) {   // one wave per Green function
        switch (h->d) {
        case 36: return launch_green_2d<P, 36, 8, 8>(h, a, g);
        case 54: return launch_green_2d<P, 54, 8, 8>(h, a, g);
        default: break;
        }
    }
    if (h->kernel_variant == 3) {   // two waves per Green function
        switch (h->d) {
        case 54: return launch_green_2d<P, 54, 8, 16>(h, a, g);
        case 72: return launch_green_2d<P, 72, 8, 16>(h, a, g);
        default: break;
        }
    }
    if (g.eigV) {
        const bool h2 = sizeof(P) == sizeof(float) && h->green_f16 && !a.raw && h->spectral_chunk > 0 && g.eigS;
        switch (h->d) {
        case 18: return launch_green_spectral<P, 18>(h, a, g);
        case 36: return launch_green_spectral<P, 36>(h, a, g);
        case 54: return launch_green_spectral<P, 54>(h, a, g);
        case 72:
            if (h2) return launch_green_spectral_tiled_h2<72, 16>(h, a, g);
            return launch_green_spectral_tiled<P, 72, 80, 1>(h, a, g);
        case 144:
            if (h2) return launch_green_spectral_tiled_h2<144, 16>(h, a, g);
            return launch_green_spectral_tiled<P, 144, 48, 1>(h, a, g);
        case 288:   // FP32: 48-column tiles (110 KB of LDS); FP64 keeps 32 (48 would take 221 KB)
            if constexpr (sizeof(P) == sizeof(float)) return launch_green_spectral_tiled<P, 288, 48, 2>(h, a, g);
            else return launch_green_spectral_tiled<P, 288, 32, 2>(h, a, g);
        default: break;
        }
    }
    if (h->d == 288) return launch_green_block<P>(h, a, g);
    switch (h->d) {
    case 18: return launch_green_2d<P, 18, 8, 8>(h, a, g);
    case 36: return launch_green_2d<P, 36, 8, 16>(h, a, g);
    case 40: return launch_green_2d<P, 40, 8, 16>(h, a, g);
    case 54: return launch_green_2d<P, 54, 16, 16>(h, a, g);
    case 72: return launch_green_2d<P, 72, 16, 16>(h, a, g);
    case 144: return launch_green_2d<P, 144, 16, 32>(h, a, g);
    default: break;
    }
    if (h->d <= 24) return launch_green_t<P, 64>(h, a, g);
    return launch_green_t<P, 256>(h, a, g);
}

template <typename P, int NT>
int launch_green_t(th_handle *h, const th_green_args &a, const green_params &g)
{
    size_t lds = sizeof(double2) * (size_t)h->d * h->d + sizeof(double2) * h->nhop +
                 2 * sizeof(double2) * h->d + 2 * sizeof(int) * h->d + 32;
    hipLaunchKernelGGL((green_kernel<P, NT>), dim3((unsigned)a.np), dim3(NT), lds, gstream(h, a), g);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

}  // namespace

int *th_status_word(th_handle *h);

// Eigenpair set the spectral kernels use for one side (nullptr: inversion)
const th_eig_set *th_side_eig(const th_handle *h, bool qside, const double *q)
{
    const bool qzero = q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0;
    if (h->spectral && h->eig_valid[0] && (!qside || qzero)) return &h->eig[0];
    if (h->spectral && qside && h->eig_valid[1] && h->eig[1].q[0] == q[0] && h->eig[1].q[1] == q[1] &&
        h->eig[1].q[2] == q[2])
        return &h->eig[1];
    return nullptr;
}

long th_green_wave_pairs(th_handle *h)
{
    if (h->wave_pairs >= 0) return h->wave_pairs;
    long q = 0;
    int dev = 0, ncu = 0, occ = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) {
        hipError_t e = hipErrorInvalidValue;
        if (h->d == 18)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, green_spectral_h2_kernel<18, TH_RUN>, 64 * 2, 0);
        else if (h->d == 36)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, green_spectral_h2_kernel<36, TH_RUN>, 64 * 3, 0);
        else if (h->d == 54)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, green_spectral_h2_kernel<54, TH_RUN>, 64 * 4, 0);
        if (e == hipSuccess && occ > 0) q = (long)TH_RUN * ncu * occ;
    }
    (void)hipGetLastError();
    h->wave_pairs = q;
    return q;
}

bool th_green_presplit_ok(const th_handle *h, const double *q)
{
    if (!(h->presplit && h->kernel_variant == 0 && h->spectral_chunk > 0 && h->green_f16)) return false;
    const th_eig_set *ea = th_side_eig(h, true, q), *eb = th_side_eig(h, false, q);
    if (!ea || !eb) return false;
    // the fp16 kernels read the split planes of V
    return (h->d == 18 || h->d == 36 || h->d == 54 || h->d == 72 || h->d == 144) && ea->S && eb->S;
}

int th_launch_green(th_handle *h, const th_green_args &a)
{
    if (h->d > TH_MAX_D_LDS && h->d != 72 && h->d != 144 && h->d != 288) {
        th_set_error("d = 2*nOrb*nA = " + std::to_string(h->d) + " has no inversion kernel");
        return TH_E_UNSUPPORTED;
    }
    green_params g;
    g.mode = a.mode; g.prec = a.prec;
    g.nA = h->nA; g.nOrb = h->nOrb; g.bs = h->bs; g.d = h->d; g.nhop = h->nhop;
    g.Mpad = h->Mpad; g.nA2 = h->nA2; g.p0 = a.p0; g.np = a.np; g.nk_re = h->nk_re;
    g.nx = a.nx; g.kmajor_re = a.kmajor_re;
    if (a.kmajor_re && a.nx < 1) {
        th_set_error("th_launch_green: k-major real pairs need the node count");
        return TH_E_ARG;
    }
    g.q0 = a.q[0]; g.q1 = a.q[1]; g.q2 = a.q[2];
    g.omega = a.omega; g.Ef = a.Ef; g.eta = a.eta; g.etap = a.etap;
    g.omega_p = a.omega_p;
    g.R = h->d_R; g.jb = h->d_jb; g.ib = h->d_ib; g.t = h->d_t; g.h0 = h->d_h0; g.hoff = h->d_hoff;
    g.kim = h->d_kim; g.yim = h->d_yim; g.wim = h->d_wim;
    g.kre = h->d_kre; g.wre = h->d_wre; g.x2 = h->d_x2; g.p2 = h->d_p2;
    g.panel = a.panel;
    g.slot0 = 0;
    g.nrun_out = (a.sum_order && h->pair_order) ? (unsigned)(a.np / TH_RUN) : 0u;
    g.presplit = a.presplit;
    g.pexp = a.pexp;
    g.pnx = a.pnx;
    g.pareal = a.pareal;
    g.raw = a.raw;
    g.status = th_status_word(h);
    // spectral Green functions when the eigenpairs of this side's k-points are ready
    const bool qside = a.mode == TH_G_IMAG_A || a.mode == TH_G_REAL_A;
    const th_eig_set *es = th_side_eig(h, qside, a.q);
    g.eigV = es ? es->V : nullptr;
    g.eigL = es ? es->L : nullptr;
    g.eigS = es ? reinterpret_cast<const _Float16 *>(es->S) : nullptr;
    g.kid_im = h->d_kid_im;
    g.kid_re = h->d_kid_re;
    g.blkmax = a.blkmax;
    h->last_blkmax = 0;
    h->n_launch[0] += 1;
    if (a.prec == TH_PREC_FP64) return launch_green_p<double>(h, a, g);
    return launch_green_p<float>(h, a, g);
}

namespace {

// H(k + q) of the unique k-points k0 .. k0 + gridDim.x - 1, column-major
// (LAPACK's layout), assembled as the Jacobi kernel assembles it
__global__ __launch_bounds__(256) void assemble_hk_kernel(green_params g, const double *klist, long k0,
                                                          const int *kidx, double q0, double q1, double q2,
                                                          double2 *Hcm)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double> *ph = reinterpret_cast<cplx<double> *>(smem);
    const int tid = threadIdx.x, d = g.d, nO = g.nOrb, nO2 = nO * nO;
    const long kb = kidx ? (long)kidx[blockIdx.x] : k0 + blockIdx.x;
    const double ka = klist[3 * kb] + q0, kbb = klist[3 * kb + 1] + q1, kc = klist[3 * kb + 2] + q2;
    for (int n = tid; n < g.nhop; n += blockDim.x) {
        const double a = ka * g.R[3 * n] + kbb * g.R[3 * n + 1] + kc * g.R[3 * n + 2];
        double sn, cs;
        sincos(a, &sn, &cs);
        ph[n] = {cs, sn};
    }
    __syncthreads();
    double2 *o = Hcm + (size_t)blockIdx.x * d * d;
    for (int e = tid; e < d * d; e += blockDim.x) {
        const int c = e / d, r = e - (e / d) * d;   // column-major: e = r + c d
        const double2 hv = g.h0[r + (size_t)c * d];
        cplx<double> x = {hv.x, hv.y};
        const int ja = r / (2 * nO), rsp = (r / nO) & 1, nu = r - (r / nO) * nO;
        const int ia = c / (2 * nO), csp = (c / nO) & 1, mu = c - (c / nO) * nO;
        if (rsp == csp) {
            const int h1 = g.hoff[ja * g.nA + ia], h2 = g.hoff[ja * g.nA + ia + 1];
            for (int n = h1; n < h2; ++n) {
                const double2 tv = g.t[(size_t)n * nO2 + nu + nO * mu];
                const cplx<double> tp = cmul(cplx<double>{tv.x, tv.y}, ph[n]);
                x.x += tp.x;
                x.y += tp.y;
            }
        }
        o[e] = make_double2(x.x, x.y);
    }
}

// V[k][r d + m] = A[k][r + m d]: eigenvectors from LAPACK's columns into the
// row-major layout of the spectral kernels (matrix k of A to slot kidx[k] of
// V where an index list is given, else to slot k)
__global__ __launch_bounds__(256) void transpose_v_kernel(const double2 *A, long n, int d, const int *kidx,
                                                          double2 *V)
{
    const size_t dd = (size_t)d * d;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)n * dd) return;
    const size_t k = e / dd;
    const int rem = (int)(e - k * dd), r = rem / d, m = rem - (rem / d) * d;
    const size_t slot = kidx ? (size_t)kidx[k] : k;
    V[slot * dd + (size_t)r * d + m] = A[k * dd + r + (size_t)m * d];
}

// flag[k] = 1 where eigenvalues or eigenvectors of matrix k are not finite
__global__ __launch_bounds__(256) void eig_finite_kernel(const double2 *A, const double *L, int d, int *flag)
{
    const size_t dd = (size_t)d * d;
    const double2 *a = A + (size_t)blockIdx.x * dd;
    const double *l = L + (size_t)blockIdx.x * d;
    int bad = 0;
    for (size_t e = threadIdx.x; e < dd; e += 256) bad |= !(isfinite(a[e].x) && isfinite(a[e].y));
    for (int e = threadIdx.x; e < d; e += 256) bad |= !isfinite(l[e]);
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) flag[blockIdx.x] = bad;
}

// Eigenpairs where d is too large for the LDS-resident Jacobi kernel: H(k + q)
// assembled on the device, rocSOLVER's batched divide-and-conquer zheevd on
// the handle's stream in chunks of at most 2 GiB of matrices, eigenvectors
// transposed into the set's layout.  zheevd has returned non-finite
// eigenvectors with info = 0 for strongly degenerate spectra (H at Gamma of
// a 16-site supercell, d = 288); such matrices are found by a finiteness
// check and decomposed again by rocSOLVER's Jacobi zheevj.  Once per sweep
// and set.
int launch_eig_dense(th_handle *h, int which, const double *q)
{
    const int d = h->d;
    const size_t dd = (size_t)d * d;
    th_eig_set &es = h->eig[which];
    if (!h->rb) {
        rocblas_handle rb = nullptr;
        if (rocblas_create_handle(&rb) != rocblas_status_success) {
            th_set_error("eigen-decomposition: rocblas_create_handle failed");
            return TH_E_HIP;
        }
        h->rb = rb;
    }
    rocblas_handle rb = static_cast<rocblas_handle>(h->rb);
    if (rocblas_set_stream(rb, h->stream) != rocblas_status_success) {
        th_set_error("eigen-decomposition: rocblas_set_stream failed");
        return TH_E_HIP;
    }
    const long chunk = std::max<long>(1, std::min<long>(h->n_kuniq, (long)(((size_t)2 << 30) / (dd * sizeof(double2)))));
    double2 *A = nullptr;
    double *E = nullptr, *W = nullptr, *res = nullptr;
    rocblas_int *info = nullptr, *nsw = nullptr;
    int *flag = nullptr, *kidx = nullptr;
    auto cleanup = [&]() {
        hipFree(A); hipFree(E); hipFree(W); hipFree(res); hipFree(info); hipFree(nsw); hipFree(flag); hipFree(kidx);
    };
    if (hipMalloc(&A, chunk * dd * sizeof(double2)) != hipSuccess || hipMalloc(&E, chunk * d * sizeof(double)) != hipSuccess ||
        hipMalloc(&W, chunk * d * sizeof(double)) != hipSuccess || hipMalloc(&res, chunk * sizeof(double)) != hipSuccess ||
        hipMalloc(&info, chunk * sizeof(rocblas_int)) != hipSuccess ||
        hipMalloc(&nsw, chunk * sizeof(rocblas_int)) != hipSuccess || hipMalloc(&flag, chunk * sizeof(int)) != hipSuccess ||
        hipMalloc(&kidx, chunk * sizeof(int)) != hipSuccess) {
        cleanup();
        (void)hipGetLastError();
        th_set_error("eigen-decomposition: out of device memory");
        return TH_E_ALLOC;
    }
    green_params g{};
    g.nA = h->nA; g.nOrb = h->nOrb; g.bs = h->bs; g.d = d; g.nhop = h->nhop;
    g.R = h->d_R; g.jb = h->d_jb; g.ib = h->d_ib; g.t = h->d_t; g.h0 = h->d_h0; g.hoff = h->d_hoff;
    std::vector<rocblas_int> hinfo(chunk);
    std::vector<int> hflag(chunk);
    const size_t lds = sizeof(double2) * std::max(1, h->nhop);
    int rc = TH_OK;
    auto fail = [&](const char *what) { th_set_error(std::string("eigen-decomposition: ") + what); return TH_E_HIP; };
    for (long k0 = 0; k0 < h->n_kuniq && !rc; k0 += chunk) {
        const long n = std::min(chunk, h->n_kuniq - k0);
        double *Lk = es.L + (size_t)k0 * d;
        hipLaunchKernelGGL(assemble_hk_kernel, dim3((unsigned)n), dim3(256), lds, h->stream, g,
                           (const double *)h->d_kuniq, k0, (const int *)nullptr, q[0], q[1], q[2], A);
        if (hipGetLastError() != hipSuccess) { rc = fail("H assembly"); break; }
        if (rocsolver_zheevd_strided_batched(rb, rocblas_evect_original, rocblas_fill_upper, d,
                                             reinterpret_cast<rocblas_double_complex *>(A), d, (rocblas_stride)dd, Lk,
                                             d, E, d, info, (rocblas_int)n) != rocblas_status_success) {
            rc = fail("rocsolver_zheevd_strided_batched failed");
            break;
        }
        hipLaunchKernelGGL(eig_finite_kernel, dim3((unsigned)n), dim3(256), 0, h->stream, (const double2 *)A,
                           (const double *)Lk, d, flag);
        hipLaunchKernelGGL(transpose_v_kernel, dim3((unsigned)((n * dd + 255) / 256)), dim3(256), 0, h->stream,
                           (const double2 *)A, n, d, (const int *)nullptr, es.V + (size_t)k0 * dd);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(hinfo.data(), info, n * sizeof(rocblas_int), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipMemcpyAsync(hflag.data(), flag, n * sizeof(int), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = fail("transpose or status copy");
            break;
        }
        // the matrices zheevd did not decompose, again by Jacobi rotations
        std::vector<int> redo;
        for (long i = 0; i < n; ++i)
            if (hinfo[i] != 0 || hflag[i]) redo.push_back((int)(k0 + i));
        if (redo.empty()) continue;
        const long nr = (long)redo.size();
        if (hipMemcpyAsync(kidx, redo.data(), nr * sizeof(int), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            rc = fail("index copy");
            break;
        }
        hipLaunchKernelGGL(assemble_hk_kernel, dim3((unsigned)nr), dim3(256), lds, h->stream, g,
                           (const double *)h->d_kuniq, 0L, (const int *)kidx, q[0], q[1], q[2], A);
        if (rocsolver_zheevj_strided_batched(rb, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, d,
                                             reinterpret_cast<rocblas_double_complex *>(A), d, (rocblas_stride)dd, 0.0,
                                             res, 100, nsw, W, d, info, (rocblas_int)nr) != rocblas_status_success) {
            rc = fail("rocsolver_zheevj_strided_batched failed");
            break;
        }
        hipLaunchKernelGGL(eig_finite_kernel, dim3((unsigned)nr), dim3(256), 0, h->stream, (const double2 *)A,
                           (const double *)W, d, flag);
        hipLaunchKernelGGL(transpose_v_kernel, dim3((unsigned)((nr * dd + 255) / 256)), dim3(256), 0, h->stream,
                           (const double2 *)A, nr, d, (const int *)kidx, es.V);
        for (long i = 0; i < nr && !rc; ++i)
            if (hipMemcpyAsync(es.L + (size_t)redo[i] * d, W + (size_t)i * d, d * sizeof(double),
                               hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
                rc = fail("eigenvalue copy");
        if (rc) break;
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(hinfo.data(), info, nr * sizeof(rocblas_int), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipMemcpyAsync(hflag.data(), flag, nr * sizeof(int), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = fail("Jacobi status copy");
            break;
        }
        for (long i = 0; i < nr; ++i)
            if (hinfo[i] != 0 || hflag[i]) {
                rc = TH_E_HIP;
                th_set_error("eigen-decomposition of H(k) failed at unique k-point " + std::to_string(redo[i]));
                break;
            }
    }
    cleanup();
    return rc;
}

}  // namespace

bool th_spectral_d(int d) { return d == 18 || d == 36 || d == 54 || d == 72 || d == 144 || d == 288; }

void th_eig_release(th_handle *h)
{
    if (h && h->rb) {
        rocblas_destroy_handle(static_cast<rocblas_handle>(h->rb));
        h->rb = nullptr;
    }
}

size_t th_eig_split_halves(int d)
{
    if (d != 18 && d != 36 && d != 54 && d != 72 && d != 144) return 0;
    return (size_t)4 * (16 * ((d + 15) / 16)) * (32 * ((d + 31) / 32));
}

namespace {

// the split planes of 2^14 V, one thread per (k-point, padded row, padded
// eigen-index): hi = fp16(x), lo = fp16(x - hi) with x the FP32 value
__global__ __launch_bounds__(256) void split_v_kernel(const double2 *V, long nk, int d, int RP, int KP,
                                                      _Float16 *S)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, pl = (size_t)RP * KP;
    if (e >= (size_t)nk * pl) return;
    const size_t k = e / pl, rm = e - k * pl;
    const int r = (int)(rm / KP), m = (int)(rm - (size_t)r * KP);
    double2 x = make_double2(0.0, 0.0);
    if (r < d && m < d) x = V[(k * d + r) * d + m];
    const float xr = (float)x.x * 16384.0f, xi = (float)x.y * 16384.0f;
    const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
    _Float16 *o = S + k * 4 * pl + rm;
    o[0] = hr;
    o[pl] = (_Float16)(xr - (float)hr);
    o[2 * pl] = hi;
    o[3 * pl] = (_Float16)(xi - (float)hi);
}

int launch_split_v(th_handle *h, const th_eig_set &es)
{
    const int d = h->d, RP = 16 * ((d + 15) / 16), KP = 32 * ((d + 31) / 32);
    const size_t n = (size_t)h->n_kuniq * RP * KP;
    hipLaunchKernelGGL(split_v_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, es.V, h->n_kuniq,
                       d, RP, KP, reinterpret_cast<_Float16 *>(es.S));
    TH_HIP(hipGetLastError());
    return TH_OK;
}

}  // namespace

// Eigenpairs of H(k + q) for the handle's unique k-points into set 'which'
int th_launch_eig(th_handle *h, int which, const double *q)
{
    green_params g{};
    g.nA = h->nA; g.nOrb = h->nOrb; g.bs = h->bs; g.d = h->d; g.nhop = h->nhop;
    g.R = h->d_R; g.jb = h->d_jb; g.ib = h->d_ib; g.t = h->d_t; g.h0 = h->d_h0; g.hoff = h->d_hoff;
    th_eig_set &es = h->eig[which];
    const size_t lds = sizeof(double2) * std::max(1, h->nhop);
    switch (h->d) {
    case 18:
        hipLaunchKernelGGL((eig_jacobi_kernel<18, 128>), dim3((unsigned)h->n_kuniq), dim3(128), lds, h->stream, g,
                           (const double *)h->d_kuniq, q[0], q[1], q[2], es.V, es.L);
        break;
    case 36:
        hipLaunchKernelGGL((eig_jacobi_kernel<36, 256>), dim3((unsigned)h->n_kuniq), dim3(256), lds, h->stream, g,
                           (const double *)h->d_kuniq, q[0], q[1], q[2], es.V, es.L);
        break;
    case 54:
        hipLaunchKernelGGL((eig_jacobi_kernel<54, 256>), dim3((unsigned)h->n_kuniq), dim3(256), lds, h->stream, g,
                           (const double *)h->d_kuniq, q[0], q[1], q[2], es.V, es.L);
        break;
    case 72:
    case 144:
    case 288: {
        const int rc = launch_eig_dense(h, which, q);
        if (rc || !es.S) return rc;
        return launch_split_v(h, es);
    }
    default:
        th_set_error("spectral Green functions are built for d = 18, 36, 54, 72, 144, 288");
        return TH_E_UNSUPPORTED;
    }
    TH_HIP(hipGetLastError());
    return es.S ? launch_split_v(h, es) : TH_OK;
}
