// th_dyson.hip -- RPA Dyson step and orbital trace on gfx950 (FP64).
//
// The reference forms T = I + chi0 U, inverts it and multiplies,
// chi = T^-1 chi0 (calculate_chi.F90:84-89).  U (Umatorb,
// mod_susceptibilities.F90:106-159) is nonzero only on the d-orbital
// diagonal-pair indices of each atom, a set S of r = O(20 nA) chi indices.
// With P the selection onto S and U_S = U[S,S], the push-through identity
//   (I + chi0 P U_S P^T)^-1 = I - chi0 P U_S (I + P^T chi0 P U_S)^-1 P^T
// gives the same chi from an r x r inverse:
//   W = I + chi0[S,S] U_S,   chi = chi0 - (chi0[:,S] U_S) (W^-1 chi0[S,:]).
// All of the near-pole ill-conditioning of (I + chi0 U) sits in W; its
// 1-norm condition number is returned.
#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

#include "th_internal.h"

// Device scratch of th_dyson_noalloc: one array per device (each GPU the
// library runs on holds its own copy until the process ends).  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device, never through the array's name.
static __device__ __align__(256) unsigned char scratch[TH_NOALLOC_SCRATCH];
static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;   // [device] -> address of scratch on that device

namespace {

__device__ inline double2 zmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double zabs1(double2 a) { return fabs(a.x) + fabs(a.y); }
__device__ inline double2 zinv(double2 a)
{
    double n = a.x * a.x + a.y * a.y;
    return make_double2(a.x / n, -a.y / n);
}

// W = I + chi0[S,S] U_S   (column-major r x r)
__global__ void form_w(const double2 *chi0, int D, const int *S, const double2 *US, int r, double2 *W)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= r * r) return;
    int a = e % r, b = e / r;
    double2 s = make_double2(a == b ? 1.0 : 0.0, 0.0);
    for (int c = 0; c < r; ++c) s = zadd(s, zmul(chi0[S[a] + (size_t)D * S[c]], US[c + (size_t)r * b]));
    W[e] = s;
}

// In-place Gauss-Jordan inverse of the column-major r x r matrix W by one
// workgroup; norms[0] = ||W||_1, norms[1] = ||W^-1||_1, status = first
// exactly-zero pivot (1-based) or 0.
__global__ __launch_bounds__(1024) void gj_global(double2 *W, int r, int *perm, double *norms, int *status)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double2 *colk = reinterpret_cast<double2 *>(smem);
    double2 *prow = colk + r;
    __shared__ double redv[1024];
    __shared__ int redi[1024];
    __shared__ int s_p;
    __shared__ double2 s_inv;
    const int tid = threadIdx.x, NT = blockDim.x;
#define Wm(a, b) W[(a) + (size_t)r * (b)]
    // ||W||_1
    double m = 0.0;
    for (int c = tid; c < r; c += NT) {
        double s = 0.0;
        for (int a = 0; a < r; ++a) s += hypot(Wm(a, c).x, Wm(a, c).y);
        m = fmax(m, s);
    }
    redv[tid] = m;
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (tid < off) redv[tid] = fmax(redv[tid], redv[tid + off]);
        __syncthreads();
    }
    if (tid == 0) norms[0] = redv[0];
    __syncthreads();
    for (int k = 0; k < r; ++k) {
        double v = -1.0;
        int ix = r;
        for (int t = k + tid; t < r; t += NT) {
            double a = zabs1(Wm(t, k));
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
        __syncthreads();   // every thread has read the pivot before row p changes
        for (int t = tid; t < r; t += NT) {
            double2 rk = Wm(k, t), rp = Wm(p, t);
            double2 ck = (t == k) ? Wm(p, k) : (t == p ? Wm(k, k) : Wm(t, k));
            colk[t] = ck;
            prow[t] = rp;        // unscaled pivot row (old row p)
            if (p != k) Wm(p, t) = rk;  // safe: columns t are owned by this thread in this phase
        }
        __syncthreads();
        if (piv.x == 0.0 && piv.y == 0.0) {
            if (tid == 0) atomicMin(status, k + 1);
            piv = make_double2(1.0, 0.0);
        }
        const double2 inv = zinv(piv);
        for (int t = tid; t < r; t += NT) {
            double2 pr = (t == k) ? inv : zmul(prow[t], inv);
            prow[t] = pr;
            Wm(k, t) = pr;
        }
        if (tid == 0) s_inv = inv;
        __syncthreads();
        for (size_t e = tid; e < (size_t)r * r; e += NT) {
            int a = (int)(e % r), c = (int)(e / r);
            if (a == k) continue;
            double2 f = colk[a];
            if (c == k) {
                double2 v = zmul(f, inv);
                Wm(a, c) = make_double2(-v.x, -v.y);
            } else {
                Wm(a, c) = zsub(Wm(a, c), zmul(f, prow[c]));
            }
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
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
    m = 0.0;
    for (int c = tid; c < r; c += NT) {
        double s = 0.0;
        for (int a = 0; a < r; ++a) s += hypot(Wm(a, c).x, Wm(a, c).y);
        m = fmax(m, s);
    }
    redv[tid] = m;
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (tid < off) redv[tid] = fmax(redv[tid], redv[tid + off]);
        __syncthreads();
    }
    if (tid == 0) norms[1] = redv[0];
#undef Wm
}

// Z = W^-1 chi0[S,:]   (r x D, column-major)
__global__ void form_z(const double2 *Winv, int r, const double2 *chi0, int D, const int *S, double2 *Z)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)r * D) return;
    int a = (int)(e % r);
    size_t y = e / r;
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < r; ++c) s = zadd(s, zmul(Winv[a + (size_t)r * c], chi0[S[c] + (size_t)D * y]));
    Z[e] = s;
}

// V = chi0[:,S] U_S   (D x r, column-major)
__global__ void form_v(const double2 *chi0, int D, const int *S, const double2 *US, int r, double2 *V)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * r) return;
    size_t x = e % D;
    int a = (int)(e / D);
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < r; ++c) s = zadd(s, zmul(chi0[x + (size_t)D * S[c]], US[c + (size_t)r * a]));
    V[e] = s;
}

// chi = chi0 - V Z
__global__ void rank_update(const double2 *chi0, const double2 *V, const double2 *Z, int D, int r, double2 *chi)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * D) return;
    size_t x = e % D, y = e / D;
    double2 s = chi0[e];
    for (int a = 0; a < r; ++a) s = zsub(s, zmul(V[x + (size_t)D * a], Z[a + (size_t)r * y]));
    chi[e] = s;
}

// schi[(sigma,i),(sigma',j)] = sum_{mu,nu} chi[(sigma,i,mu,mu),(sigma',j,nu,nu)]
__global__ void trace_kernel(const double2 *chi, int nA, int nOrb, int D, double2 *schi)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    int n4 = 4 * nA;
    if (e >= n4 * n4) return;
    int row = e % n4, col = e / n4;
    int sg = row / nA, i = row % nA, sp = col / nA, j = col % nA;
    double2 s = make_double2(0.0, 0.0);
    for (int nu = 0; nu < nOrb; ++nu)
        for (int mu = 0; mu < nOrb; ++mu) {
            int X = 4 * ((i * nOrb + mu) * nOrb + mu) + sg;
            int Y = 4 * ((j * nOrb + nu) * nOrb + nu) + sp;
            s = zadd(s, chi[X + (size_t)D * Y]);
        }
    schi[e] = s;
}

inline unsigned nblk(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

// writes ||W||_1 = ||W^-1||_1 = 1 (U = 0: chi = chi0, condition number 1)
__global__ void unit_norms(double *norms)
{
    if (threadIdx.x < 2) norms[threadIdx.x] = 1.0;
}

}  // namespace

int th_dyson_prepare(th_handle *h, const double2 *U, int *r_out)
{
    const int D = h->D;
    // support of U: indices with a nonzero row or column
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
    *r_out = r;
    if (!h->d_X) TH_HIP(hipMalloc(&h->d_X, sizeof(double) * 2 + sizeof(int) * (size_t)std::max(r, 1)));
    if (r == 0) return TH_OK;
    std::vector<double2> US((size_t)r * r);
    for (int b = 0; b < r; ++b)
        for (int a = 0; a < r; ++a) US[a + (size_t)r * b] = U[S[a] + (size_t)D * S[b]];
    if (r > h->sup_cap) {
        hipFree(h->d_sup); hipFree(h->d_U); hipFree(h->d_W); hipFree(h->d_V); hipFree(h->d_Z); hipFree(h->d_X);
        h->d_sup = nullptr; h->d_U = h->d_W = h->d_V = h->d_Z = h->d_X = nullptr;
        h->sup_cap = 0;
        TH_HIP(hipMalloc(&h->d_sup, sizeof(int) * (r + 2)));
        TH_HIP(hipMalloc(&h->d_U, sizeof(double2) * (size_t)r * r));
        TH_HIP(hipMalloc(&h->d_W, sizeof(double2) * (size_t)r * r));
        TH_HIP(hipMalloc(&h->d_V, sizeof(double2) * (size_t)D * r));
        TH_HIP(hipMalloc(&h->d_Z, sizeof(double2) * (size_t)r * D));
        TH_HIP(hipMalloc(&h->d_X, sizeof(double) * 2 + sizeof(int) * (size_t)r));
        h->sup_cap = r;
    }
    TH_HIP(hipMemcpyAsync(h->d_sup, S.data(), sizeof(int) * r, hipMemcpyHostToDevice, h->stream));
    TH_HIP(hipMemcpyAsync(h->d_U, US.data(), sizeof(double2) * (size_t)r * r, hipMemcpyHostToDevice, h->stream));
    // the host vectors go out of scope: the copies must have landed
    TH_HIP(hipStreamSynchronize(h->stream));
    return TH_OK;
}

int th_dyson_launch(th_handle *h, int r, const double2 *chi0, double2 *chi, double *norms_dev)
{
    const int D = h->D;
    double *norms = norms_dev ? norms_dev : reinterpret_cast<double *>(h->d_X);
    if (r == 0) {
        TH_HIP(hipMemcpyAsync(chi, chi0, sizeof(double2) * (size_t)D * D, hipMemcpyDeviceToDevice, h->stream));
        hipLaunchKernelGGL(unit_norms, dim3(1), dim3(64), 0, h->stream, norms);
        TH_HIP(hipGetLastError());
        return TH_OK;
    }
    int *perm = reinterpret_cast<int *>(reinterpret_cast<double *>(h->d_X) + 2);
    hipLaunchKernelGGL(form_w, dim3(nblk((size_t)r * r, 256)), dim3(256), 0, h->stream, chi0, D, h->d_sup, h->d_U,
                       r, h->d_W);
    TH_HIP(hipGetLastError());
    hipLaunchKernelGGL(gj_global, dim3(1), dim3(1024), sizeof(double2) * 2 * r, h->stream, h->d_W, r, perm, norms,
                       th_status_word(h));
    TH_HIP(hipGetLastError());
    hipLaunchKernelGGL(form_z, dim3(nblk((size_t)r * D, 256)), dim3(256), 0, h->stream, h->d_W, r, chi0, D,
                       h->d_sup, h->d_Z);
    TH_HIP(hipGetLastError());
    hipLaunchKernelGGL(form_v, dim3(nblk((size_t)r * D, 256)), dim3(256), 0, h->stream, chi0, D, h->d_sup, h->d_U,
                       r, h->d_V);
    TH_HIP(hipGetLastError());
    hipLaunchKernelGGL(rank_update, dim3(nblk((size_t)D * D, 256)), dim3(256), 0, h->stream, chi0, h->d_V, h->d_Z,
                       D, r, chi);
    TH_HIP(hipGetLastError());
    h->n_launch[3] += 5;
    return TH_OK;
}

int th_dyson_device(th_handle *h, const double2 *chi0, const double2 *U, double2 *chi, double *cond)
{
    int r = 0;
    int rc = th_dyson_prepare(h, U, &r);
    if (!rc) rc = th_dyson_launch(h, r, chi0, chi, nullptr);
    if (rc) return rc;
    if (cond) {
        double nv[2];
        TH_HIP(hipMemcpyAsync(nv, h->d_X, sizeof(nv), hipMemcpyDeviceToHost, h->stream));
        TH_HIP(hipStreamSynchronize(h->stream));
        *cond = nv[0] * nv[1];
    }
    return TH_OK;
}

int th_launch_trace(th_handle *h, const double2 *chi, double2 *schi)
{
    int n4 = 4 * h->nA;
    hipLaunchKernelGGL(trace_kernel, dim3(nblk((size_t)n4 * n4, 64)), dim3(64), 0, h->stream, chi, h->nA, h->nOrb,
                       h->D, schi);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

int th_noalloc_init(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return TH_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        th_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return TH_E_HIP;
    }
    g_scratch_ptr[device] = p;
    return TH_OK;
}

namespace {

size_t up256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

int th_dyson_noalloc_impl(th_handle *h, const double2 *chi0, const double2 *U, double2 *chi, double *cond)
{
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (h->device < (int)g_scratch_ptr.size()) base = static_cast<unsigned char *>(g_scratch_ptr[h->device]);
    }
    if (!base) {
        th_set_error("th_dyson_noalloc: the device scratch was not set up by th_create on this device");
        return TH_E_STATE;
    }
    const int D = h->D;
    std::vector<char> on(D, 0);
    for (int c = 0; c < D; ++c)
        for (int a = 0; a < D; ++a) {
            const double2 u = U[a + (size_t)D * c];
            if (u.x != 0.0 || u.y != 0.0) { on[a] = 1; on[c] = 1; }
        }
    std::vector<int> S;
    for (int a = 0; a < D; ++a)
        if (on[a]) S.push_back(a);
    const int r = (int)S.size();
    // scratch layout: norms[2] + status | perm[r] | S[r] | U_S | W | V | Z
    const size_t o_perm = 256, o_S = o_perm + up256(sizeof(int) * (size_t)std::max(r, 1));
    const size_t o_US = o_S + up256(sizeof(int) * (size_t)std::max(r, 1));
    const size_t o_W = o_US + up256(sizeof(double2) * (size_t)r * r);
    const size_t o_V = o_W + up256(sizeof(double2) * (size_t)r * r);
    const size_t o_Z = o_V + up256(sizeof(double2) * (size_t)D * r);
    const size_t need = o_Z + up256(sizeof(double2) * (size_t)r * D);
    if (need > TH_NOALLOC_SCRATCH) {
        th_set_error("th_dyson_noalloc: support of U (" + std::to_string(r) + " indices) with D = " +
                     std::to_string(D) + " needs " + std::to_string(need) + " bytes of scratch, the limit is " +
                     std::to_string((size_t)TH_NOALLOC_SCRATCH));
        return TH_E_UNSUPPORTED;
    }
    double *norms = reinterpret_cast<double *>(base);
    int *status = reinterpret_cast<int *>(base + 2 * sizeof(double));
    int *perm = reinterpret_cast<int *>(base + o_perm);
    int *dS = reinterpret_cast<int *>(base + o_S);
    double2 *dUS = reinterpret_cast<double2 *>(base + o_US), *W = reinterpret_cast<double2 *>(base + o_W);
    double2 *V = reinterpret_cast<double2 *>(base + o_V), *Z = reinterpret_cast<double2 *>(base + o_Z);
    // the scratch keeps the previous call's contents: the status word starts
    // at 'no zero pivot'
    const int none = INT_MAX;
    TH_HIP(hipMemcpyAsync(status, &none, sizeof(int), hipMemcpyHostToDevice, h->stream));
    if (r == 0) {
        TH_HIP(hipMemcpyAsync(chi, chi0, sizeof(double2) * (size_t)D * D, hipMemcpyDeviceToDevice, h->stream));
        hipLaunchKernelGGL(unit_norms, dim3(1), dim3(64), 0, h->stream, norms);
        TH_HIP(hipGetLastError());
    } else {
        std::vector<double2> US((size_t)r * r);
        for (int b = 0; b < r; ++b)
            for (int a = 0; a < r; ++a) US[a + (size_t)r * b] = U[S[a] + (size_t)D * S[b]];
        TH_HIP(hipMemcpyAsync(dS, S.data(), sizeof(int) * r, hipMemcpyHostToDevice, h->stream));
        TH_HIP(hipMemcpyAsync(dUS, US.data(), sizeof(double2) * (size_t)r * r, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(form_w, dim3(nblk((size_t)r * r, 256)), dim3(256), 0, h->stream, chi0, D, dS, dUS, r, W);
        TH_HIP(hipGetLastError());
        hipLaunchKernelGGL(gj_global, dim3(1), dim3(1024), sizeof(double2) * 2 * r, h->stream, W, r, perm, norms,
                           status);
        TH_HIP(hipGetLastError());
        hipLaunchKernelGGL(form_z, dim3(nblk((size_t)r * D, 256)), dim3(256), 0, h->stream, W, r, chi0, D, dS, Z);
        TH_HIP(hipGetLastError());
        hipLaunchKernelGGL(form_v, dim3(nblk((size_t)r * D, 256)), dim3(256), 0, h->stream, chi0, D, dS, dUS, r, V);
        TH_HIP(hipGetLastError());
        hipLaunchKernelGGL(rank_update, dim3(nblk((size_t)D * D, 256)), dim3(256), 0, h->stream, chi0, V, Z, D, r,
                           chi);
        TH_HIP(hipGetLastError());
    }
    double nv[2];
    int st = INT_MAX;
    TH_HIP(hipMemcpyAsync(nv, norms, sizeof(nv), hipMemcpyDeviceToHost, h->stream));
    TH_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    // also keeps the host copies of S and U_S alive until their uploads are done
    TH_HIP(hipStreamSynchronize(h->stream));
    if (cond) *cond = nv[0] * nv[1];
    if (st != INT_MAX) {
        th_set_error("th_dyson_noalloc: exactly singular pivot of (I + chi0 U), step " + std::to_string(st));
        return st;
    }
    return TH_OK;
}
