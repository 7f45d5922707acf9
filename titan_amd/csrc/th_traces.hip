// Orbital traces of one susceptibility on the device: the spin traces, the
// spin-orbit cross responses <<L,S>>, <<S,L>>, <<L,L>> in cartesian
// components, and the rotation of the spin traces to the local frames
// (the post-processing of calculate_chi.F90:91-185).
//
// chi is D x D complex column-major (chi[X + D Y]) with the reference's
// index X = 4 ((i nOrb + mu) nOrb + nu) + sigma, sigma = (+, up, down, -);
// traces are indexed s nA + i and stored column-major.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "th_internal.h"

namespace {

__device__ inline double2 zmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// spin (+, up, down, -) <-> cartesian (0, x, y, z), [row * 4 + col]
// (mod_constants.F90:34-45: StoC[c][s], CtoS[s][c])
__constant__ double2 c_stoc[16] = {
    {0.0, 0.0}, {1.0, 0.0}, {1.0, 0.0}, {0.0, 0.0},
    {0.5, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.5, 0.0},
    {0.0, -0.5}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.5},
    {0.0, 0.0}, {0.5, 0.0}, {-0.5, 0.0}, {0.0, 0.0}};
__constant__ double2 c_ctos[16] = {
    {0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}, {0.0, 0.0},
    {0.5, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0},
    {0.5, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {-1.0, 0.0},
    {0.0, 0.0}, {1.0, 0.0}, {0.0, -1.0}, {0.0, 0.0}};

__device__ inline int cx(int nOrb, int s, int i, int mu, int nu) { return 4 * ((i * nOrb + mu) * nOrb + nu) + s; }

// spin traces schi[(s,i),(t,j)] = sum_{mu,nu} chi[(s,i,mu,mu),(t,j,nu,nu)]
__global__ void spin_trace_kernel(const double2 *chi, int nA, int nOrb, int D, double2 *schi)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x, n4 = 4 * nA;
    if (e >= n4 * n4) return;
    const int row = e % n4, col = e / n4;
    const int s = row / nA, i = row % nA, t = col / nA, j = col % nA;
    double2 acc = make_double2(0.0, 0.0);
    for (int nu = 0; nu < nOrb; ++nu)
        for (int mu = 0; mu < nOrb; ++mu)
            acc = zadd(acc, chi[cx(nOrb, s, i, mu, mu) + (size_t)D * cx(nOrb, t, j, nu, nu)]);
    schi[e] = acc;
}

// cross responses (calculate_chi.F90:109-140), one output entry per thread:
// kind 0 LS, 1 SL, 2 LL; lvec[mu + nOrb (gama + nOrb s)] = <mu|L_s|gama>
__global__ void soc_trace_kernel(const double2 *chi, int nA, int nOrb, int D, const double2 *lvec, double2 *ls,
                                 double2 *sl, double2 *ll)
{
    const int n3 = 3 * nA, per = n3 * n3;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3 * per) return;
    const int kind = e / per, f = e - kind * per;
    const int row = f % n3, col = f / n3;
    const int s = row / nA, i = row % nA, t = col / nA, j = col % nA;
    auto L = [&](int mu, int ga, int c) { return lvec[mu + nOrb * (ga + nOrb * c)]; };
    auto X = [&](int a, int b) { return chi[a + (size_t)D * b]; };
    double2 acc = make_double2(0.0, 0.0);
    if (kind == 0) {          // LS: L(mu,ga,s) (chi[(up|dn,i,mu,ga),(p,j,nu,nu)]) CtoS(p, t+1)
        for (int mu = 0; mu < nOrb; ++mu)
            for (int ga = 0; ga < nOrb; ++ga) {
                const double2 l = L(mu, ga, s);
                if (l.x == 0.0 && l.y == 0.0) continue;
                double2 r = make_double2(0.0, 0.0);
                for (int nu = 0; nu < nOrb; ++nu)
                    for (int p = 0; p < 4; ++p) {
                        const int Y = cx(nOrb, p, j, nu, nu);
                        const double2 v = zadd(X(cx(nOrb, 1, i, mu, ga), Y), X(cx(nOrb, 2, i, mu, ga), Y));
                        r = zadd(r, zmul(v, c_ctos[p * 4 + t + 1]));
                    }
                acc = zadd(acc, zmul(l, r));
            }
        ls[f] = acc;
    } else if (kind == 1) {   // SL: StoC(s+1, p) (chi[(p,i,mu,mu),(up|dn,j,nu,xi)]) L(nu,xi,t)
        for (int nu = 0; nu < nOrb; ++nu)
            for (int xi = 0; xi < nOrb; ++xi) {
                const double2 l = L(nu, xi, t);
                if (l.x == 0.0 && l.y == 0.0) continue;
                double2 r = make_double2(0.0, 0.0);
                for (int mu = 0; mu < nOrb; ++mu)
                    for (int p = 0; p < 4; ++p) {
                        const int Xi = cx(nOrb, p, i, mu, mu);
                        const double2 v = zadd(X(Xi, cx(nOrb, 1, j, nu, xi)), X(Xi, cx(nOrb, 2, j, nu, xi)));
                        r = zadd(r, zmul(c_stoc[(s + 1) * 4 + p], v));
                    }
                acc = zadd(acc, zmul(r, l));
            }
        sl[f] = acc;
    } else {                  // LL: L(mu,ga,s) (chi[(up|dn),(up|dn)]) L(nu,xi,t)
        for (int mu = 0; mu < nOrb; ++mu)
            for (int ga = 0; ga < nOrb; ++ga) {
                const double2 l = L(mu, ga, s);
                if (l.x == 0.0 && l.y == 0.0) continue;
                double2 r = make_double2(0.0, 0.0);
                for (int nu = 0; nu < nOrb; ++nu)
                    for (int xi = 0; xi < nOrb; ++xi) {
                        const double2 m = L(nu, xi, t);
                        if (m.x == 0.0 && m.y == 0.0) continue;
                        double2 v = make_double2(0.0, 0.0);
                        for (int a = 1; a <= 2; ++a)
                            for (int b = 1; b <= 2; ++b) v = zadd(v, X(cx(nOrb, a, i, mu, ga), cx(nOrb, b, j, nu, xi)));
                        r = zadd(r, zmul(v, m));
                    }
                acc = zadd(acc, zmul(l, r));
            }
        ll[f] = acc;
    }
}

// block (i, j) of the spin traces -> A_i block B_j; rot[(2 i + {0: A, 1: B}) 16 + r 4 + c]
__global__ void rotate_kernel(const double2 *schi, int nA, const double2 *rot, double2 *out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x, n4 = 4 * nA;
    if (e >= n4 * n4) return;
    const int row = e % n4, col = e / n4;
    const int s = row / nA, i = row % nA, t = col / nA, j = col % nA;
    const double2 *A = rot + (size_t)(2 * i) * 16, *B = rot + (size_t)(2 * j + 1) * 16;
    double2 acc = make_double2(0.0, 0.0);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            acc = zadd(acc, zmul(zmul(A[s * 4 + a], schi[(a * nA + i) + (size_t)n4 * (b * nA + j)]), B[b * 4 + t]));
    out[e] = acc;
}

inline unsigned nblk(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

// left (A) or right (B) rotation of a spin-spin susceptibility in the
// (+, up, down, -) basis, angles in degrees (mod_rotation_matrices.F90:6-76)
void rotation_pair(double theta_deg, double phi_deg, double2 *A, double2 *B)
{
    const double d2r = std::acos(-1.0) / 180.0, t = theta_deg * d2r, p = phi_deg * d2r;
    const double a = 1.0 + std::cos(t), b = 1.0 - std::cos(t), s = std::sin(t);
    const double2 e = make_double2(std::cos(p), std::sin(p)), em = make_double2(std::cos(p), -std::sin(p));
    auto sc = [](double2 z, double f) { return make_double2(0.5 * z.x * f, 0.5 * z.y * f); };
    auto re = [](double f) { return make_double2(0.5 * f, 0.0); };
    const double2 Av[16] = {sc(em, a), re(-s), re(s), sc(e, -b),
                            sc(em, s), re(a), re(b), sc(e, s),
                            sc(em, -s), re(b), re(a), sc(e, -s),
                            sc(em, -b), re(-s), re(s), sc(e, a)};
    const double2 Bv[16] = {sc(e, a), sc(e, s), sc(e, -s), sc(e, -b),
                            re(-s), re(a), re(b), re(-s),
                            re(s), re(b), re(a), re(s),
                            sc(em, -b), sc(em, s), sc(em, -s), sc(em, a)};
    for (int k = 0; k < 16; ++k) { A[k] = Av[k]; B[k] = Bv[k]; }
}

}  // namespace

extern "C" int th_chi_traces(th_handle *h, const double *chi, const double *lvec, const double *angles, double *schi,
                             double *ls, double *sl, double *ll)
{
    if (!h || (!schi && !ls && !sl && !ll) || (lvec == nullptr && (ls || sl || ll)) || (angles && !schi)) {
        th_set_error("th_chi_traces: bad arguments (outputs, or lvec for the cross responses, or schi for angles)");
        return TH_E_ARG;
    }
    const int nA = h->nA, nOrb = h->nOrb, D = h->D, n4 = 4 * nA, n3 = 3 * nA;
    const size_t DD = (size_t)D * D;
    double2 *d_src = nullptr, *d_tmp = nullptr, *d_l = nullptr, *d_rot = nullptr, *d_out = nullptr;
    auto cleanup = [&]() { hipFree(d_src); hipFree(d_tmp); hipFree(d_l); hipFree(d_rot); hipFree(d_out); };
    const double2 *src = h->d_chi;
    if (chi) {
        if (hipMalloc(&d_src, sizeof(double2) * DD) != hipSuccess) { cleanup(); th_set_error("th_chi_traces: out of device memory"); return TH_E_HIP; }
        TH_HIP(hipMemcpyAsync(d_src, chi, sizeof(double2) * DD, hipMemcpyHostToDevice, h->stream));
        src = d_src;
    } else if (!src) {
        th_set_error("th_chi_traces: no chi given and none computed on this handle");
        return TH_E_STATE;
    }
    const size_t nout = (size_t)n4 * n4 * 2 + (size_t)3 * n3 * n3;
    if (hipMalloc(&d_out, sizeof(double2) * nout) != hipSuccess) { cleanup(); th_set_error("th_chi_traces: out of device memory"); return TH_E_HIP; }
    double2 *d_schi = d_out, *d_rs = d_out + (size_t)n4 * n4, *d_x = d_rs + (size_t)n4 * n4;
    if (schi) {
        hipLaunchKernelGGL(spin_trace_kernel, dim3(nblk((size_t)n4 * n4, 64)), dim3(64), 0, h->stream, src, nA, nOrb, D,
                           d_schi);
        if (angles) {
            std::vector<double2> rot((size_t)nA * 32);
            for (int i = 0; i < nA; ++i)
                rotation_pair(180.0 - angles[2 * i], 180.0 + angles[2 * i + 1], &rot[(size_t)32 * i], &rot[(size_t)32 * i + 16]);
            if (hipMalloc(&d_rot, sizeof(double2) * rot.size()) != hipSuccess) { cleanup(); th_set_error("th_chi_traces: out of device memory"); return TH_E_HIP; }
            TH_HIP(hipMemcpyAsync(d_rot, rot.data(), sizeof(double2) * rot.size(), hipMemcpyHostToDevice, h->stream));
            hipLaunchKernelGGL(rotate_kernel, dim3(nblk((size_t)n4 * n4, 64)), dim3(64), 0, h->stream, d_schi, nA, d_rot,
                               d_rs);
            TH_HIP(hipMemcpyAsync(schi, d_rs, sizeof(double2) * n4 * n4, hipMemcpyDeviceToHost, h->stream));
        } else {
            TH_HIP(hipMemcpyAsync(schi, d_schi, sizeof(double2) * n4 * n4, hipMemcpyDeviceToHost, h->stream));
        }
    }
    if (lvec) {
        if (hipMalloc(&d_l, sizeof(double2) * (size_t)nOrb * nOrb * 3) != hipSuccess) { cleanup(); th_set_error("th_chi_traces: out of device memory"); return TH_E_HIP; }
        TH_HIP(hipMemcpyAsync(d_l, lvec, sizeof(double2) * (size_t)nOrb * nOrb * 3, hipMemcpyHostToDevice, h->stream));
        const size_t per = (size_t)n3 * n3;
        hipLaunchKernelGGL(soc_trace_kernel, dim3(nblk(3 * per, 64)), dim3(64), 0, h->stream, src, nA, nOrb, D, d_l,
                           d_x, d_x + per, d_x + 2 * per);
        if (ls) TH_HIP(hipMemcpyAsync(ls, d_x, sizeof(double2) * per, hipMemcpyDeviceToHost, h->stream));
        if (sl) TH_HIP(hipMemcpyAsync(sl, d_x + per, sizeof(double2) * per, hipMemcpyDeviceToHost, h->stream));
        if (ll) TH_HIP(hipMemcpyAsync(ll, d_x + 2 * per, sizeof(double2) * per, hipMemcpyDeviceToHost, h->stream));
    }
    TH_HIP(hipGetLastError());
    TH_HIP(hipStreamSynchronize(h->stream));
    cleanup();
    return TH_OK;
}
