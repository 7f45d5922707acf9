// th_refine.hip -- chi0 on the support of U in FP64, from the eigenpairs of H(k).
//
// The Dyson step (calculate_chi.F90:84-89) through the support S of U is
//   chi = chi0 - chi0[:,S] U_S W^-1 chi0[S,:],   W = I + chi0[S,S] U_S.
// Near the Goldstone/FMR pole W is nearly singular: an error of chi0[S,S]
// reaches chi multiplied by the condition number of W, while errors of every
// other entry of chi0 enter chi linearly.  The mixed-precision chi0 therefore
// gets its S x S block recomputed in FP64 here, from the eigen-decompositions
// H(k) = V diag(l) V^H that its spectral Green functions already use.
//
// With G = V diag(f) V^H, f_n = 1/(z - l_n), each term of eintshechi.F90:86-104
// (imaginary axis) and :146-164 (real axis) is
//   chi0[X,Y] += c_mn Vq[x,m] conj(V[x',n]) conj(Vq[y,m]) V[y',n]   (summed over m, n)
// where X = (sigma, i, mu, nu) selects the G_q row x = (i, a_sigma, nu) and the
// G column x' = (i, b_sigma, mu), and Y = (sigma', j, gamma, xi) the G_q column
// y = (j, a_sigma', gamma) and the G row y' = (j, b_sigma', xi), with (a, b) =
// (dn, up), (up, up), (dn, dn), (up, dn) for sigma = +, up, dn, -, and
//   imaginary axis  c_mn = w (fq_m f_n + conj(fq_m f_n)) = 2 w Re(fq_m f_n)
//   real axis       c_mn = -i w (fq_m - conj fq_m) conj(f_n) = 2 w Im(fq_m) conj(f_n).
// The energy sum is taken first, K_mn(k) = sum_E c_mn(E) for each k-point, and
//   chi0[X,Y] = (1/2pi) sum_k sum_n conj(V[x',n]) V[y',n] C_n[x,y],
//   C_n[x,y]  = sum_m Vq[x,m] K_mn conj(Vq[y,m]),
// |R|^2 d^2 + |S|^2 d complex products per k-point (R: the rows x, x' that S
// uses) where a per-pair evaluation of the same block costs 2 |R|^2 d per
// (k,E) pair.  Each workgroup sums its k-points in a fixed order into its own
// partial block; one kernel adds the partials in a fixed order and writes the
// block into chi0 (bitwise reproducible, no atomics).
#include <algorithm>
#include <cmath>

#include "th_internal.h"

namespace {

constexpr int RF_NT = 512;     // threads per workgroup
constexpr int RF_ACC = 8;      // accumulator entries (X, Y) per thread
constexpr int RF_EC = 16;      // energies per LDS chunk of the K sum
constexpr size_t RF_LDS = 160 * 1024;

struct refine_params {
    int nR, nS, nx_chunk;
    const int *R;              // [nR] Hamiltonian rows
    const int *sx, *sxp;       // [nS] index in R of x_X and of x'_X (S sorted by x_X)
    const int *sy, *syp;       // [nS] index in R of y_Y and of y'_Y
    const int *tiles;          // [ntile][4]: S rows s0..s1, R range r0..r1 of their x
    long nk;
    int kpc;                   // k-points per chunk
    const double2 *Vq, *V;     // [nk][d][d] row-major: V[r][m] = component r of eigenvector m
    const double *Lq, *L;      // [nk][d]
    const int *im_ptr, *im_idx;   // imaginary pairs of each k-point (CSR over stored pairs)
    const double *yim, *wim;
    const int *re_ptr, *re_idx;   // real-mesh points of each k-point (CSR)
    const double *wre, *x2, *p2;
    int nx, real_axis;
    double omega, Ef, eta, etap;
    const double *omega_p;
    double2 *part;             // [nchunk][nS][nS] column-major blocks
};

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a conj(b)
__device__ inline double2 cmulc(double2 a, double2 b) { return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__device__ inline double2 cfma(double2 a, double2 b, double2 c)
{
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
// 1 / (re + i im)
__device__ inline double2 crecip(double re, double im)
{
    const double n = re * re + im * im;
    return make_double2(re / n, -im / n);
}

// Loads the eigenvector rows of R and the eigenvalues of k-point k into LDS and
// sums K_mn over the k-point's energies into sK (row pitch PITCH); sW is the
// scratch of one energy chunk.  Ends with K stored, without a barrier.
template <int D, int PITCH>
__device__ void load_and_k(const refine_params &p, long k, double omega, double2 *sVq, double2 *sV, double2 *sK,
                           double *sLq, double *sL, double2 *sW)
{
    constexpr int NKE = (D * D + RF_NT - 1) / RF_NT;
    const int nR = p.nR, tid = threadIdx.x;
    // eigenvector rows of R and the eigenvalues
    const double2 *Vqk = p.Vq + (size_t)k * D * D, *Vk = p.V + (size_t)k * D * D;
    for (int e = tid; e < nR * D; e += RF_NT) {
        const int r = e / D, m = e - r * D;
        sVq[e] = Vqk[(size_t)p.R[r] * D + m];
        sV[e] = Vk[(size_t)p.R[r] * D + m];
    }
    for (int m = tid; m < D; m += RF_NT) {
        sLq[m] = p.Lq[(size_t)k * D + m];
        sL[m] = p.L[(size_t)k * D + m];
    }
    // K_mn = sum over the energies of this k-point
    double kr[NKE], ki[NKE];
#pragma unroll
    for (int j = 0; j < NKE; ++j) kr[j] = ki[j] = 0.0;
    const int i0 = p.im_ptr[k], i1 = p.im_ptr[k + 1];
    const int q0 = p.real_axis ? p.re_ptr[k] : 0, q1 = p.real_axis ? p.re_ptr[k + 1] : 0;
    const long nim = i1 - i0, ntot = nim + (long)(q1 - q0) * p.nx;
    double2 *sA = sW, *sB = sW + RF_EC * D;
    for (long c0 = 0, ne = 0; c0 < ntot; c0 += ne) {
        // chunks never mix the two axes: imaginary energies first
        const bool imag = c0 < nim;
        ne = imag ? min((long)RF_EC, nim - c0) : min((long)RF_EC, ntot - c0);
        __syncthreads();   // previous chunk consumed (first pass: V, L stored)
        for (int v = tid; v < 2 * ne * D; v += RF_NT) {
            const int side = v / (ne * D), e = (v - side * ne * D) / D, m = v - side * ne * D - e * D;
            double2 val;
            if (imag) {
                const int pr = p.im_idx[i0 + c0 + e];
                const double y = p.yim[pr], w = p.wim[pr];
                if (side == 0) {
                    const double2 f = crecip(p.Ef + omega - sLq[m], y + p.eta);
                    val = make_double2(2.0 * w * f.x, 2.0 * w * f.y);
                } else {
                    val = crecip(p.Ef - sL[m], y + p.etap);
                }
            } else {
                const long g = c0 - nim + e;
                const int rm = p.re_idx[q0 + (int)(g / p.nx)], node = (int)(g % p.nx);
                const double w = p.p2[node] * p.wre[rm], x = p.x2[node];
                if (side == 0) {
                    val = make_double2(2.0 * w * crecip(x + omega - sLq[m], p.eta).y, 0.0);
                } else {
                    const double2 f = crecip(x - sL[m], p.etap);
                    val = make_double2(f.x, -f.y);
                }
            }
            (side ? sB : sA)[e * D + m] = val;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NKE; ++j) {
            const int e = tid + RF_NT * j;
            if (e < D * D) {
                const int m = e / D, n = e - m * D;
                double a = kr[j], b = ki[j];
                if (imag) {
                    for (int t = 0; t < ne; ++t) {
                        const double2 x = sA[t * D + m], y = sB[t * D + n];
                        a = fma(x.x, y.x, fma(-x.y, y.y, a));
                    }
                } else {
                    for (int t = 0; t < ne; ++t) {
                        const double x = sA[t * D + m].x;
                        const double2 y = sB[t * D + n];
                        a = fma(x, y.x, a);
                        b = fma(x, y.y, b);
                    }
                }
                kr[j] = a;
                ki[j] = b;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NKE; ++j) {
        const int e = tid + RF_NT * j;
        if (e < D * D) {
            const int m = e / D, n = e - m * D;
            sK[m * PITCH + n] = make_double2(kr[j], ki[j]);
        }
    }
}

template <int D>
__global__ __launch_bounds__(RF_NT) void refine_kernel(refine_params p)
{
    constexpr int DP = (D + 3) / 4 * 4;     // padded row of K and C
    constexpr int NQ = DP / 4;              // column quads of a C row
    extern __shared__ __align__(16) unsigned char smem[];
    const int nR = p.nR, nS = p.nS, tid = threadIdx.x;
    double2 *sVq = reinterpret_cast<double2 *>(smem);      // [nR][D]
    double2 *sV = sVq + nR * D;                            // [nR][D]
    double2 *sK = sV + nR * D;                             // [D][DP]
    double *sLq = reinterpret_cast<double *>(sK + D * DP); // [D]
    double *sL = sLq + D;                                  // [D]
    double2 *sW = reinterpret_cast<double2 *>(sL + D);     // energy chunk [2][RF_EC][D] or C [nx_chunk][nR][DP]
    const int *tl = p.tiles + 4 * blockIdx.y;
    const int s0 = tl[0], s1 = tl[1], r0 = tl[2], r1 = tl[3];
    const int rows = s1 - s0;
    const double omega = p.omega_p ? *p.omega_p : p.omega;
    double2 acc[RF_ACC];
    int ex[RF_ACC], ey[RF_ACC];
#pragma unroll
    for (int j = 0; j < RF_ACC; ++j) {
        acc[j] = make_double2(0.0, 0.0);
        const int e = tid + RF_NT * j;
        ex[j] = e < rows * nS ? s0 + e % rows : -1;
        ey[j] = e < rows * nS ? e / rows : 0;
    }
    const long k0 = (long)blockIdx.x * p.kpc, k1 = min(p.nk, k0 + p.kpc);
    for (long k = k0; k < k1; ++k) {
        load_and_k<D, DP>(p, k, omega, sVq, sV, sK, sLq, sL, sW);
        if constexpr (DP > D)
            for (int e = tid; e < D * (DP - D); e += RF_NT)
                sK[(e / (DP - D)) * DP + D + e % (DP - D)] = make_double2(0.0, 0.0);
        // C_n[x,y] for the tile's x, a chunk of x values at a time, and the
        // S x S entries whose x lies in the chunk
        for (int xa = r0; xa < r1; xa += p.nx_chunk) {
            const int nxc = min(p.nx_chunk, r1 - xa);
            __syncthreads();   // K stored / previous C consumed
            const int nyp = (nR + 1) / 2, ntl = nxc * nyp * NQ;
            for (int t = tid; t < ntl; t += RF_NT) {
                const int q = t % NQ, rest = t / NQ, yp = rest % nyp, xl = rest / nyp;
                const int ya = 2 * yp, yb = min(2 * yp + 1, nR - 1);
                const double2 *vx = sVq + (xa + xl) * D, *va = sVq + ya * D, *vb = sVq + yb * D;
                double2 c[2][4];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) c[i][j] = make_double2(0.0, 0.0);
#pragma unroll 2
                for (int m = 0; m < D; ++m) {
                    const double2 x = vx[m];
                    const double2 a0 = cmulc(x, va[m]), a1 = cmulc(x, vb[m]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double2 kk = sK[m * DP + q + NQ * j];
                        c[0][j] = cfma(a0, kk, c[0][j]);
                        c[1][j] = cfma(a1, kk, c[1][j]);
                    }
                }
                double2 *cx = sW + (size_t)xl * nR * DP;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cx[ya * DP + q + NQ * j] = c[0][j];
                    if (yb != ya) cx[yb * DP + q + NQ * j] = c[1][j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < RF_ACC; ++j) {
                const int X = ex[j];
                if (X < 0) continue;
                const int xr = p.sx[X];
                if (xr < xa || xr >= xa + nxc) continue;
                const int Y = ey[j];
                const double2 *cr = sW + ((size_t)(xr - xa) * nR + p.sy[Y]) * DP;
                const double2 *vxp = sV + p.sxp[X] * D, *vyp = sV + p.syp[Y] * D;
                double2 s = acc[j];
#pragma unroll 2
                for (int n = 0; n < D; ++n) s = cfma(cr[n], cmulc(vyp[n], vxp[n]), s);
                acc[j] = s;
            }
        }
        __syncthreads();   // C consumed before the next k-point's loads
    }
    double2 *out = p.part + (size_t)blockIdx.x * nS * nS;
#pragma unroll
    for (int j = 0; j < RF_ACC; ++j)
        if (ex[j] >= 0) out[ex[j] + (size_t)nS * ey[j]] = acc[j];
}

typedef double dv4 __attribute__((ext_vector_type(4)));

// Supports of at most 32 indices on the FP64 matrix cores: per k-point the block
// is one GEMM over the d^2 eigen-index pairs (m, n),
//   chi0_SS(k) = A B,   A[X, (m,n)] = Vq[x_X, m] conj(V[x'_X, n]) K_mn,
//                       B[(m,n), Y] = conj(Vq[y_Y, m]) V[y'_Y, n],
// 4 pairs per v_mfma_f64_16x16x4 step, the operands formed in registers from
// the eigenvector rows in LDS and the complex product
// as 4 real MFMAs.  Each of the 8 waves takes every 8th step and keeps its own
// 32 x 32 accumulators over the workgroup's k-points; the waves' sums are
// added in a fixed order through LDS at the end.  Lane l supplies A[l & 15][l >> 4]
// and B[l >> 4][l & 15] and holds C[(l >> 4) + 4 q][l & 15] in register q.
template <int D>
__global__ __launch_bounds__(RF_NT) void refine_mfma_kernel(refine_params p)
{
    constexpr int NSTEP = D * D / 4;
    static_assert(D * D % 4 == 0, "pairs (m, n) must fill whole MFMA steps");
    extern __shared__ __align__(16) unsigned char smem[];
    const int nR = p.nR, nS = p.nS, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double2 *sVq = reinterpret_cast<double2 *>(smem);       // [nR][D]
    double2 *sV = sVq + nR * D;                             // [nR][D]
    double2 *sK = sV + nR * D;                              // [D][D]
    double *sLq = reinterpret_cast<double *>(sK + D * D);   // [D]
    double *sL = sLq + D;                                   // [D]
    double2 *sW = reinterpret_cast<double2 *>(sL + D);      // energy chunk; at the end the 32 x 32 sums
    const double omega = p.omega_p ? *p.omega_p : p.omega;
    const int nt = (nS + 15) >> 4;                          // 16-wide tiles per dimension (1 or 2)
    // this lane's rows of A (X = 16 t + (lane & 15): x, x') and columns of B
    // (Y = 16 t + (lane & 15): y, y') as row offsets in sVq / sV
    int px[2], ppx[2], py[2], ppy[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int X = 16 * t + (lane & 15);
        px[t] = X < nS ? p.sx[X] * D : -1;
        ppx[t] = X < nS ? p.sxp[X] * D : -1;
        py[t] = X < nS ? p.sy[X] * D : -1;
        ppy[t] = X < nS ? p.syp[X] * D : -1;
    }
    dv4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) { cr[a][b][q] = 0.0; ci[a][b][q] = 0.0; }
    const long k0 = (long)blockIdx.x * p.kpc, k1 = min(p.nk, k0 + p.kpc);
    for (long k = k0; k < k1; ++k) {
        load_and_k<D, D>(p, k, omega, sVq, sV, sK, sLq, sL, sW);
        __syncthreads();   // K stored
        for (int st = wave; st < NSTEP; st += RF_NT / 64) {
            const int idx = 4 * st + (lane >> 4), m = idx / D, n = idx - m * D;
            const double2 kk = sK[idx];
            double ar[2], ai[2], br[2], bi[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
                if (t < nt && px[t] >= 0) {
                    a = cmul(cmulc(sVq[px[t] + m], sV[ppx[t] + n]), kk);
                    b = cmulc(sV[ppy[t] + n], sVq[py[t] + m]);
                }
                ar[t] = a.x; ai[t] = a.y; br[t] = b.x; bi[t] = b.y;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (a < nt && b < nt) {
                        cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
                        cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[b], cr[a][b], 0, 0, 0);
                        ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
                        ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
                    }
                }
        }
        __syncthreads();   // this k-point's rows and K consumed
    }
    // waves' sums added one wave at a time in wave order: sum[X][Y] in LDS
    double2 *sum = sW;   // [32][32]
    for (int w = 0; w < RF_NT / 64; ++w) {
        if (wave == w) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int X = 16 * a + (lane >> 4) + 4 * q, Y = 16 * b + (lane & 15);
                        const double2 v = make_double2(cr[a][b][q], ci[a][b][q]);
                        sum[X * 32 + Y] = w ? make_double2(sum[X * 32 + Y].x + v.x, sum[X * 32 + Y].y + v.y) : v;
                    }
        }
        __syncthreads();
    }
    double2 *out = p.part + (size_t)blockIdx.x * nS * nS;
    for (int e = tid; e < nS * nS; e += RF_NT) {
        const int X = e % nS, Y = e / nS;
        out[e] = sum[X * 32 + Y];
    }
}

// chi0[S[a], S[b]] = (1/2pi) sum_c part[c][a][b], chunks in order
__global__ void refine_reduce_kernel(const double2 *__restrict__ part, int nchunk, int nS, const int *__restrict__ S,
                                     int Dchi, double2 *__restrict__ chi0)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nS * nS) return;
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < nchunk; ++c) {
        const double2 v = part[(size_t)c * nS * nS + e];
        s.x += v.x;
        s.y += v.y;
    }
    const double inv2pi = 1.0 / (2.0 * 3.14159265358979323846);
    const int a = e % nS, b = e / nS;
    chi0[S[a] + (size_t)Dchi * S[b]] = make_double2(s.x * inv2pi, s.y * inv2pi);
}

template <int D> int refine_smem_fixed(int nR) { return (2 * nR * D + D * ((D + 3) / 4 * 4)) * 16 + 2 * D * 8; }
template <int D> int refine_dp() { return (D + 3) / 4 * 4; }

}  // namespace

// Support of U and everything derived from it (host side; allocations here,
// none in the launch).
int th_refine_set_support(th_handle *h, const double2 *U)
{
    h->ref_S.clear();
    h->ref_ready = 0;
    if (!U) return TH_OK;
    const int D = h->D;
    std::vector<char> on(D, 0);
    for (int c = 0; c < D; ++c)
        for (int a = 0; a < D; ++a) {
            const double2 u = U[a + (size_t)D * c];
            if (u.x != 0.0 || u.y != 0.0) { on[a] = 1; on[c] = 1; }
        }
    for (int a = 0; a < D; ++a)
        if (on[a]) h->ref_S.push_back(a);
    return TH_OK;
}

namespace {

// row of the Hamiltonian for atom i, spin s (0 up, 1 down), orbital o
inline int hrow(const th_handle *h, int i, int s, int o) { return (2 * i + s) * h->nOrb + o; }

int refine_build(th_handle *h)
{
    const int nS = (int)h->ref_S.size();
    const int d = h->d, nO = h->nOrb;
    // sigma = +, up, dn, -: G_q row spin a, G column spin b
    const int sa[4] = {1, 0, 1, 0}, sb[4] = {0, 0, 1, 1};
    // as a row X: x = (i, a, nu), x' = (i, b, mu); as a column Y: y = (i, a, mu),
    // y' = (i, b, nu)
    std::vector<int> x(nS), xp(nS), y(nS), yp(nS);
    for (int s = 0; s < nS; ++s) {
        const int X = h->ref_S[s], sg = X % 4, rest = X / 4, nu = rest % nO, mu = (rest / nO) % nO, i = rest / (nO * nO);
        x[s] = hrow(h, i, sa[sg], nu);
        xp[s] = hrow(h, i, sb[sg], mu);
        y[s] = hrow(h, i, sa[sg], mu);
        yp[s] = hrow(h, i, sb[sg], nu);
    }
    std::vector<int> R(x);
    R.insert(R.end(), xp.begin(), xp.end());
    R.insert(R.end(), y.begin(), y.end());
    R.insert(R.end(), yp.begin(), yp.end());
    std::sort(R.begin(), R.end());
    R.erase(std::unique(R.begin(), R.end()), R.end());
    const int nR = (int)R.size();
    std::vector<int> ridx(d, -1);
    for (int r = 0; r < nR; ++r) ridx[R[r]] = r;
    // S sorted by the R index of x (stable)
    std::vector<int> ord(nS);
    for (int s = 0; s < nS; ++s) ord[s] = s;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ridx[x[a]] < ridx[x[b]]; });
    std::vector<int> Ss(nS), sx(nS), sxp(nS), sy(nS), syp(nS);
    for (int s = 0; s < nS; ++s) {
        Ss[s] = h->ref_S[ord[s]];
        sx[s] = ridx[x[ord[s]]];
        sxp[s] = ridx[xp[ord[s]]];
        sy[s] = ridx[y[ord[s]]];
        syp[s] = ridx[yp[ord[s]]];
    }
    // tiles of S rows: at most RF_NT * RF_ACC accumulator entries each
    const int rows = std::max(1, std::min(nS, RF_NT * RF_ACC / nS));
    std::vector<int> tiles;
    for (int s0 = 0; s0 < nS; s0 += rows) {
        const int s1 = std::min(nS, s0 + rows);
        tiles.insert(tiles.end(), {s0, s1, sx[s0], sx[s1 - 1] + 1});
    }
    int ntile = (int)tiles.size() / 4;
    // supports of at most 32 indices: one GEMM per k-point on the FP64 matrix cores
    h->ref_mfma = nS <= 32 ? 1 : 0;
    if (h->ref_mfma) {
        const long fixed = (2L * nR * d + (long)d * d) * 16 + 2L * d * 8;
        h->ref_smem = (size_t)(fixed + std::max(2L * RF_EC * d * 16, 32L * 32 * 16));
        if (h->ref_smem > RF_LDS) h->ref_mfma = 0;
    }
    if (h->ref_mfma) {
        ntile = 1;
        h->ref_nxc = 1;
        h->ref_nR = nR;
        h->ref_ntile = 1;
        tiles.assign({0, nS, 0, nR});
    }
    // LDS: chunk of x values of C that fits next to V, K and the eigenvalues
    int dp = 0, fixed = 0;
    switch (d) {
    case 18: dp = refine_dp<18>(); fixed = refine_smem_fixed<18>(nR); break;
    case 36: dp = refine_dp<36>(); fixed = refine_smem_fixed<36>(nR); break;
    case 54: dp = refine_dp<54>(); fixed = refine_smem_fixed<54>(nR); break;
    default: return TH_E_UNSUPPORTED;
    }
    if (!h->ref_mfma) {
        const long avail = (long)RF_LDS - fixed;
        const long per_x = (long)nR * dp * 16, ebuf = 2L * RF_EC * d * 16;
        if (avail < std::max(per_x, ebuf)) return TH_E_UNSUPPORTED;
        const int nxc = (int)std::max(1L, std::min((long)nR, avail / per_x));
        h->ref_smem = (size_t)fixed + (size_t)std::max(ebuf, (long)nxc * per_x);
        h->ref_nxc = nxc;
        h->ref_nR = nR;
        h->ref_ntile = ntile;
    }
    // k-point chunks: enough workgroups for the chip, partials within 256 MB
    const long nk = h->n_kuniq;
    const size_t blk = (size_t)nS * nS * sizeof(double2);
    long nchunk = std::max(1L, std::min(nk, h->ref_mfma ? 1024L : std::max(1L, 4096L / ntile)));
    nchunk = std::max(1L, std::min(nchunk, (long)((256u << 20) / blk)));
    h->ref_kpc = (int)((nk + nchunk - 1) / nchunk);
    h->ref_nchunk = (int)((nk + h->ref_kpc - 1) / h->ref_kpc);
    // device copies
    std::vector<int> ints;
    ints.insert(ints.end(), R.begin(), R.end());
    ints.insert(ints.end(), sx.begin(), sx.end());
    ints.insert(ints.end(), sxp.begin(), sxp.end());
    ints.insert(ints.end(), sy.begin(), sy.end());
    ints.insert(ints.end(), syp.begin(), syp.end());
    ints.insert(ints.end(), tiles.begin(), tiles.end());
    ints.insert(ints.end(), Ss.begin(), Ss.end());
    hipFree(h->d_ref_int);
    h->d_ref_int = nullptr;
    TH_HIP(hipMalloc(&h->d_ref_int, sizeof(int) * ints.size()));
    TH_HIP(hipMemcpy(h->d_ref_int, ints.data(), sizeof(int) * ints.size(), hipMemcpyHostToDevice));
    const size_t pbytes = (size_t)h->ref_nchunk * blk;
    if (pbytes > h->ref_part_bytes) {
        hipFree(h->d_ref_part);
        h->d_ref_part = nullptr;
        h->ref_part_bytes = 0;
        TH_HIP(hipMalloc(&h->d_ref_part, pbytes));
        h->ref_part_bytes = pbytes;
    }
    h->ref_built_S = h->ref_S;
    h->ref_built_nk = nk;
    return TH_OK;
}

}  // namespace

// Host preparation before a capturable launch: true in *on when the block
// will be recomputed (mixed precision, a support set, eigenpairs of both
// sides valid, d with a kernel, support small enough).
int th_refine_prepare(th_handle *h, const double *q, int prec, bool *on)
{
    *on = false;
    const bool want = (prec == TH_PREC_MIXED && h->refine) || (prec == TH_PREC_FP64 && h->refine == 2);
    if (!want || h->ref_S.empty() || (int)h->ref_S.size() > TH_REFINE_MAX_S) return TH_OK;
    if (h->d != 18 && h->d != 36 && h->d != 54) return TH_OK;
    const bool qz = q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0;
    if (!h->spectral || !h->eig_valid[0] || (!qz && !h->eig_valid[1]) || !h->d_kim_ptr) return TH_OK;
    if (h->ref_built_S != h->ref_S || h->ref_built_nk != h->n_kuniq || !h->d_ref_part) {
        const int rc = refine_build(h);
        if (rc == TH_E_UNSUPPORTED) return TH_OK;
        if (rc) return rc;
    }
    *on = true;
    return TH_OK;
}

int th_refine_launch(th_handle *h, const double *q, double omega, const double *omega_p, double Ef, double eta,
                     double etap, int n_x, bool real_axis, double2 *chi0)
{
    const bool qz = q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0;
    const int nS = (int)h->ref_built_S.size(), nR = h->ref_nR;
    refine_params p{};
    p.nR = nR;
    p.nS = nS;
    p.nx_chunk = h->ref_nxc;
    p.R = h->d_ref_int;
    p.sx = p.R + nR;
    p.sxp = p.sx + nS;
    p.sy = p.sxp + nS;
    p.syp = p.sy + nS;
    p.tiles = p.syp + nS;
    const int *Ssorted = p.tiles + 4 * h->ref_ntile;
    p.nk = h->n_kuniq;
    p.kpc = h->ref_kpc;
    const th_eig_set &e0 = h->eig[0], &e1 = h->eig[qz ? 0 : 1];
    p.V = e0.V; p.L = e0.L;
    p.Vq = e1.V; p.Lq = e1.L;
    p.im_ptr = h->d_kim_ptr; p.im_idx = h->d_kim_idx; p.yim = h->d_yim; p.wim = h->d_wim;
    p.re_ptr = h->d_kre_ptr; p.re_idx = h->d_kre_idx; p.wre = h->d_wre;
    p.x2 = h->d_x2; p.p2 = h->d_p2;
    p.nx = real_axis ? n_x : 0;
    p.real_axis = real_axis && h->d_kre_ptr ? 1 : 0;
    p.omega = omega; p.omega_p = omega_p; p.Ef = Ef; p.eta = eta; p.etap = etap;
    p.part = reinterpret_cast<double2 *>(h->d_ref_part);
    const dim3 grid((unsigned)h->ref_nchunk, (unsigned)h->ref_ntile);
    const void *fn = nullptr;
    const bool mf = h->ref_mfma != 0;
    switch (h->d) {
    case 18: fn = mf ? reinterpret_cast<const void *>(refine_mfma_kernel<18>) : reinterpret_cast<const void *>(refine_kernel<18>); break;
    case 36: fn = mf ? reinterpret_cast<const void *>(refine_mfma_kernel<36>) : reinterpret_cast<const void *>(refine_kernel<36>); break;
    case 54: fn = mf ? reinterpret_cast<const void *>(refine_mfma_kernel<54>) : reinterpret_cast<const void *>(refine_kernel<54>); break;
    default: th_set_error("refinement built for d = 18, 36, 54"); return TH_E_UNSUPPORTED;
    }
    TH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->ref_smem));
    switch (h->d * 2 + (mf ? 1 : 0)) {
    case 36: hipLaunchKernelGGL(refine_kernel<18>, grid, dim3(RF_NT), h->ref_smem, h->stream, p); break;
    case 37: hipLaunchKernelGGL(refine_mfma_kernel<18>, grid, dim3(RF_NT), h->ref_smem, h->stream, p); break;
    case 72: hipLaunchKernelGGL(refine_kernel<36>, grid, dim3(RF_NT), h->ref_smem, h->stream, p); break;
    case 73: hipLaunchKernelGGL(refine_mfma_kernel<36>, grid, dim3(RF_NT), h->ref_smem, h->stream, p); break;
    case 108: hipLaunchKernelGGL(refine_kernel<54>, grid, dim3(RF_NT), h->ref_smem, h->stream, p); break;
    default: hipLaunchKernelGGL(refine_mfma_kernel<54>, grid, dim3(RF_NT), h->ref_smem, h->stream, p); break;
    }
    TH_HIP(hipGetLastError());
    hipLaunchKernelGGL(refine_reduce_kernel, dim3((unsigned)((nS * nS + 255) / 256)), dim3(256), 0, h->stream,
                       reinterpret_cast<const double2 *>(h->d_ref_part), h->ref_nchunk, nS, Ssorted, h->D, chi0);
    TH_HIP(hipGetLastError());
    h->n_launch[2] += 2;
    return TH_OK;
}
