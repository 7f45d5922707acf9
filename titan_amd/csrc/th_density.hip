// Local spin density <S^munu_ii> = <c+_imu sigma c_inu> by an imaginary-axis
// sweep of G (calcSmunu, mod_sumrule.F90:162-245): for every pair of the
// handle's imaginary list, G(k, E_F + i(y + eta)) on the device (the Green
// function kernels with their raw output), the four local spin blocks of
// every atom accumulated as w (G[nu s', mu s] + conj(G[mu s, nu s'])), summed
// in a fixed order (per chunk of pairs, then over chunks, then over batches).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "th_internal.h"

namespace {

constexpr int TH_SD_CHUNK = 256;   // pairs per partial sum

// partial[chunk][out], out = ((i nO + mu) nO + nu) 4 + blk, blk = uu, dd, ud, du
__global__ void density_partial_kernel(const double2 *G, long np, int nA, int nO, const double *w, long p0,
                                       double2 *partial)
{
    const int nout = nA * nO * nO * 4;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const long chunk = blockIdx.y;
    if (o >= nout) return;
    const int blk = o & 3, r = o >> 2;
    const int nu = r % nO, mu = (r / nO) % nO, i = r / (nO * nO);
    const int d = 2 * nO * nA;
    // spin of the row of G[nu, mu] (s1) and of its column (s2): uu (0,0), dd (1,1), ud (0,1), du (1,0)
    const int s1 = (blk == 1 || blk == 3) ? 1 : 0, s2 = (blk == 1 || blk == 2) ? 1 : 0;
    const int rn = (2 * i + s1) * nO + nu, cm = (2 * i + s2) * nO + mu;   // G[nu s1, mu s2]
    const int rm = (2 * i + s2) * nO + mu, cn = (2 * i + s1) * nO + nu;   // conj(G[mu s2, nu s1])
    double2 acc = make_double2(0.0, 0.0);
    const long b = chunk * TH_SD_CHUNK, e = b + TH_SD_CHUNK < np ? b + TH_SD_CHUNK : np;
    for (long p = b; p < e; ++p) {
        const double2 *g = G + (size_t)p * d * d;
        const double2 x = g[(size_t)rn * d + cm], y = g[(size_t)rm * d + cn];
        const double ww = w[p0 + p];
        acc.x += (x.x + y.x) * ww;
        acc.y += (x.y - y.y) * ww;
    }
    partial[(size_t)chunk * nout + o] = acc;
}

// acc[out] += sum over chunks of partial (fixed order)
__global__ void density_sum_kernel(const double2 *partial, long nchunk, int nout, double2 *acc)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    double2 s = acc[o];
    for (long c = 0; c < nchunk; ++c) {
        const double2 v = partial[(size_t)c * nout + o];
        s.x += v.x;
        s.y += v.y;
    }
    acc[o] = s;
}

// acc[((i nO + mu) nO + nu) 4 + blk] = sum_p w_p (G[nu s1, mu s2] + conj(G[mu s2, nu s1]))
// over the imaginary list at z = E_F + i(y_p + eta), blk = uu, dd, ud, du
int density_blocks(th_handle *h, double Ef, double eta, std::vector<double2> &acc)
{
    if (!h->d_kim || h->n_im <= 0) { th_set_error("th_set_imag_points must be called before the ground-state sweep"); return TH_E_STATE; }
    const int nA = h->nA, nO = h->nOrb, d = h->d, nout = nA * nO * nO * 4;
    // pairs per batch: the raw G (d^2 complex per pair) within the panel budget
    const size_t per = (size_t)d * d * sizeof(double2);
    const long bp = std::max<long>(TH_SD_CHUNK, std::min<long>(h->n_im, (long)(((size_t)h->batch_mb << 20) / per)));
    const long nch_max = (bp + TH_SD_CHUNK - 1) / TH_SD_CHUNK;
    double2 *d_raw = nullptr, *d_part = nullptr, *d_acc = nullptr;
    auto cleanup = [&]() { hipFree(d_raw); hipFree(d_part); hipFree(d_acc); };
    if (hipMalloc(&d_raw, per * bp) != hipSuccess || hipMalloc(&d_part, sizeof(double2) * nch_max * nout) != hipSuccess ||
        hipMalloc(&d_acc, sizeof(double2) * nout) != hipSuccess) {
        cleanup();
        (void)hipGetLastError();
        th_set_error("ground-state sweep: out of device memory");
        return TH_E_HIP;
    }
    int rc = th_ensure_panels(h, bp);
    if (!rc) rc = th_reset_status(h);
    if (!rc) rc = th_launch_zero(h, reinterpret_cast<double *>(d_acc), (size_t)2 * nout);
    for (long p0 = 0; p0 < h->n_im && !rc; p0 += bp) {
        const long np = std::min(bp, h->n_im - p0);
        th_green_args a{};
        a.mode = TH_G_IMAG_B;     // G(k, E_F + i(y + eta')) with eta' = eta
        a.p0 = p0;
        a.np = np;
        a.Ef = Ef;
        a.eta = eta;
        a.etap = eta;
        a.prec = TH_PREC_FP64;
        a.panel = h->d_pB;
        a.raw = d_raw;
        if ((rc = th_launch_green(h, a))) break;
        const long nch = (np + TH_SD_CHUNK - 1) / TH_SD_CHUNK;
        hipLaunchKernelGGL(density_partial_kernel, dim3((unsigned)((nout + 63) / 64), (unsigned)nch), dim3(64), 0,
                           h->stream, (const double2 *)d_raw, np, nA, nO, h->d_wim, p0, d_part);
        hipLaunchKernelGGL(density_sum_kernel, dim3((unsigned)((nout + 63) / 64)), dim3(64), 0, h->stream,
                           (const double2 *)d_part, nch, nout, d_acc);
        if (hipGetLastError() != hipSuccess) { rc = TH_E_HIP; th_set_error("ground-state sweep: kernel launch"); }
    }
    acc.assign(nout, make_double2(0.0, 0.0));
    if (!rc && hipMemcpyAsync(acc.data(), d_acc, sizeof(double2) * nout, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
        rc = TH_E_HIP;
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = TH_E_HIP;
    if (!rc) rc = th_check_status(h);
    cleanup();
    return rc;
}

}  // namespace

extern "C" int th_spin_density(th_handle *h, double Ef, double eta, double *S_out)
{
    if (!h || !S_out || !(eta >= 0.0)) { th_set_error("th_spin_density: bad arguments"); return TH_E_ARG; }
    std::vector<double2> acc;
    const int rc = density_blocks(h, Ef, eta, acc);
    if (rc) return rc;
    const int nA = h->nA, nO = h->nOrb;
    // / 2 pi, + 1/2 on the diagonal of uu and dd (mod_sumrule.F90:224-239), then
    // S_x = ud + du, S_y = i (ud - du), S_z = uu - dd
    const double inv2pi = 1.0 / (2.0 * std::acos(-1.0));
    for (int i = 0; i < nA; ++i)
        for (int mu = 0; mu < nO; ++mu)
            for (int nu = 0; nu < nO; ++nu) {
                const int r = ((i * nO + mu) * nO + nu) * 4;
                double2 uu = acc[r], dd = acc[r + 1], ud = acc[r + 2], du = acc[r + 3];
                uu.x *= inv2pi; uu.y *= inv2pi; dd.x *= inv2pi; dd.y *= inv2pi;
                ud.x *= inv2pi; ud.y *= inv2pi; du.x *= inv2pi; du.y *= inv2pi;
                if (mu == nu) { uu.x += 0.5; dd.x += 0.5; }
                const double2 sx = make_double2(ud.x + du.x, ud.y + du.y);
                const double2 sy = make_double2(-(ud.y - du.y), ud.x - du.x);
                const double2 sz = make_double2(uu.x - dd.x, uu.y - dd.y);
                const double2 sv[3] = {sx, sy, sz};
                for (int k = 0; k < 3; ++k) {
                    const size_t o = (size_t)mu + nO * ((size_t)nu + nO * ((size_t)k + 3 * i));
                    S_out[2 * o] = sv[k].x;
                    S_out[2 * o + 1] = sv[k].y;
                }
            }
    return TH_OK;
}

extern "C" int th_expectation_values(th_handle *h, double Ef, double eta, double *rho, double *mx, double *my,
                                     double *mz, double *mp)
{
    if (!h || !rho || !mx || !my || !mz || !mp || !(eta >= 0.0)) {
        th_set_error("th_expectation_values: bad arguments");
        return TH_E_ARG;
    }
    std::vector<double2> acc;
    const int rc = density_blocks(h, Ef, eta, acc);
    if (rc) return rc;
    // the diagonal mu = nu of the blocks (expectation_values_greenfunction,
    // mod_expectation.F90:101-142): uu and dd hold 2 w Re G, du holds
    // w (G[d,u] + conj G[u,d]); / pi, + 1/2 on uu and dd
    const int nA = h->nA, nO = h->nOrb;
    const double inv2pi = 1.0 / (2.0 * std::acos(-1.0));
    for (int i = 0; i < nA; ++i)
        for (int mu = 0; mu < nO; ++mu) {
            const int r = ((i * nO + mu) * nO + mu) * 4;
            const double uu = 0.5 + acc[r].x * inv2pi, dd = 0.5 + acc[r + 1].x * inv2pi;
            const double2 du = acc[r + 3];
            const size_t o = (size_t)mu + (size_t)nO * i;
            rho[o] = uu + dd;
            mz[o] = uu - dd;
            mx[o] = 2.0 * du.x * inv2pi;
            my[o] = 2.0 * du.y * inv2pi;
            mp[2 * o] = mx[o];
            mp[2 * o + 1] = my[o];
        }
    return TH_OK;
}
