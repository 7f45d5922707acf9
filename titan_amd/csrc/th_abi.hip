// th_abi.hip -- the C ABI (include/titan_hip.h): handle life cycle, device
// buffers and the per-frequency orchestration of the gfx950 kernels.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include "th_internal.h"

namespace {
thread_local std::string g_err = "";

// Gauss-Legendre nodes on [x1,x2] (Newton on P_n from the Tricomi guess),
// the rule of EnergyIntegration.F90:125-157, for the real-axis nodes of
// th_chi_sweep (EnergyIntegration.F90:99-122).
void gauleg(double x1, double x2, int n, double *x, double *w)
{
    const int m = (n + 1) / 2;
    const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; ++i) {
        double z = std::cos(3.141592654 * (i - 0.25) / (n + 0.5)), z1 = 0.0, pp = 0.0;
        do {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (std::fabs(z - z1) > 3.0e-14);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = (2.0 * xl) / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

struct status_block {
    int *d = nullptr;
};

// exact k-point identity (bit patterns of the three coordinates)
struct kkey {
    uint64_t a, b, c;
    bool operator==(const kkey &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct kkey_hash {
    size_t operator()(const kkey &k) const
    {
        uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
        h ^= k.b + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        h ^= k.c + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
kkey make_key(const double *k)
{
    kkey r;
    std::memcpy(&r.a, k, 8);
    std::memcpy(&r.b, k + 1, 8);
    std::memcpy(&r.c, k + 2, 8);
    return r;
}

}  // namespace

void th_set_error(const std::string &msg) { g_err = msg; }

// per-handle device status word (first zero pivot, INT_MAX if none); kept
// in a side table so the handle layout stays plain
static std::vector<std::pair<th_handle *, int *>> g_status;
static std::mutex g_status_mu;   // handles may be created and used from several threads

int *th_status_word(th_handle *h)
{
    std::lock_guard<std::mutex> lk(g_status_mu);
    for (auto &e : g_status)
        if (e.first == h) return e.second;
    return nullptr;
}

int th_reset_status(th_handle *h)
{
    const int v = INT_MAX;
    TH_HIP(hipMemcpyAsync(th_status_word(h), &v, sizeof(int), hipMemcpyHostToDevice, h->stream));
    return TH_OK;
}

int th_check_status(th_handle *h)
{
    int v = INT_MAX;
    TH_HIP(hipMemcpyAsync(&v, th_status_word(h), sizeof(int), hipMemcpyDeviceToHost, h->stream));
    TH_HIP(hipStreamSynchronize(h->stream));
    if (v != INT_MAX) {
        th_set_error("exactly singular pivot in an inversion, step " + std::to_string(v));
        return v;
    }
    return TH_OK;
}

namespace {

size_t prec_size(int prec) { return prec == TH_PREC_FP64 ? sizeof(double) : sizeof(float); }

int ensure_panels(th_handle *h, long pairs)
{
    const size_t need = (size_t)pairs * h->nA2 * 2 * h->Mpad * sizeof(double);
    if (need <= h->panel_bytes) return TH_OK;
    TH_HIP(hipStreamSynchronize(h->stream2));
    hipFree(h->d_pA);
    hipFree(h->d_pB);
    hipFree(h->d_pA2);
    hipFree(h->d_pB2);
    h->d_pA = h->d_pB = h->d_pA2 = h->d_pB2 = nullptr;
    h->panel_bytes = 0;
    TH_HIP(hipMalloc(&h->d_pA, need));
    TH_HIP(hipMalloc(&h->d_pB, need));
    TH_HIP(hipMalloc(&h->d_pA2, need));
    TH_HIP(hipMalloc(&h->d_pB2, need));
    h->panel_bytes = need;
    return TH_OK;
}

}  // namespace

int th_ensure_panels(th_handle *h, long pairs) { return ensure_panels(h, pairs); }

// partial-sum buffer of the contraction for at least 'splits' splits
int th_ensure_part(th_handle *h, int splits)
{
    if (splits <= h->part_cap) return TH_OK;
    TH_HIP(hipStreamSynchronize(h->stream));
    TH_HIP(hipStreamSynchronize(h->stream2));
    hipFree(h->d_part);
    h->d_part = nullptr;
    h->part_cap = 0;
    const size_t MM = (size_t)h->Mpad * h->Mpad;
    TH_HIP(hipMalloc(&h->d_part, sizeof(double) * (size_t)splits * h->nA2 * 2 * MM));
    h->part_cap = splits;
    return TH_OK;
}

// real-axis energy nodes of the current omega to the device
int th_upload_nodes(th_handle *h, int n_x, const double *x2, const double *p2)
{
    if (n_x > h->nx_cap) {
        hipFree(h->d_x2);
        hipFree(h->d_p2);
        h->d_x2 = h->d_p2 = nullptr;
        h->nx_cap = 0;
        TH_HIP(hipMalloc(&h->d_x2, sizeof(double) * n_x));
        TH_HIP(hipMalloc(&h->d_p2, sizeof(double) * n_x));
        h->nx_cap = n_x;
    }
    TH_HIP(hipMemcpyAsync(h->d_x2, x2, sizeof(double) * n_x, hipMemcpyHostToDevice, h->stream));
    TH_HIP(hipMemcpyAsync(h->d_p2, p2, sizeof(double) * n_x, hipMemcpyHostToDevice, h->stream));
    return TH_OK;
}

// Unique k-points of the imaginary list and the real mesh (the eigenpairs
// of H are computed once per unique k) and the maps to them.
static int rebuild_kunion(th_handle *h)
{
    ++h->list_version;
    h->eig_valid[0] = h->eig_valid[1] = 0;
    std::unordered_map<kkey, int, kkey_hash> ids;
    std::vector<double> ku;
    auto add = [&](const double *k) {
        auto it = ids.find(make_key(k));
        if (it != ids.end()) return it->second;
        const int id = (int)(ku.size() / 3);
        ids.emplace(make_key(k), id);
        ku.insert(ku.end(), k, k + 3);
        return id;
    };
    std::vector<int> map_im(h->kim_uniq.size() / 3), kid_re(h->kre_host.size() / 3);
    for (size_t u = 0; u < map_im.size(); ++u) map_im[u] = add(&h->kim_uniq[3 * u]);
    for (size_t u = 0; u < kid_re.size(); ++u) kid_re[u] = add(&h->kre_host[3 * u]);
    std::vector<int> kid_im(h->kid_im_local.size());
    for (size_t p = 0; p < kid_im.size(); ++p) kid_im[p] = map_im[h->kid_im_local[p]];
    hipFree(h->d_kuniq); hipFree(h->d_kid_im); hipFree(h->d_kid_re);
    h->d_kuniq = nullptr; h->d_kid_im = h->d_kid_re = nullptr;
    h->n_kuniq = (long)(ku.size() / 3);
    if (h->n_kuniq > 0) {
        TH_HIP(hipMalloc(&h->d_kuniq, sizeof(double) * ku.size()));
        TH_HIP(hipMemcpy(h->d_kuniq, ku.data(), sizeof(double) * ku.size(), hipMemcpyHostToDevice));
    }
    if (!kid_im.empty()) {
        TH_HIP(hipMalloc(&h->d_kid_im, sizeof(int) * kid_im.size()));
        TH_HIP(hipMemcpy(h->d_kid_im, kid_im.data(), sizeof(int) * kid_im.size(), hipMemcpyHostToDevice));
    }
    if (!kid_re.empty()) {
        TH_HIP(hipMalloc(&h->d_kid_re, sizeof(int) * kid_re.size()));
        TH_HIP(hipMemcpy(h->d_kid_re, kid_re.data(), sizeof(int) * kid_re.size(), hipMemcpyHostToDevice));
    }
    h->kid_im_host = kid_im;
    h->kid_re_host = kid_re;
    // imaginary pairs (stored order) and real-mesh points of each unique k-point
    hipFree(h->d_kim_ptr); hipFree(h->d_kim_idx); hipFree(h->d_kre_ptr); hipFree(h->d_kre_idx);
    h->d_kim_ptr = h->d_kim_idx = h->d_kre_ptr = h->d_kre_idx = nullptr;
    if (h->n_kuniq > 0) {
        auto csr = [&](const std::vector<int> &kid, int **dptr, int **didx) -> int {
            std::vector<int> ptr(h->n_kuniq + 1, 0), idx(std::max<size_t>(kid.size(), 1), 0);
            for (int u : kid) ++ptr[u + 1];
            for (long u = 0; u < h->n_kuniq; ++u) ptr[u + 1] += ptr[u];
            std::vector<int> fill(ptr.begin(), ptr.end() - 1);
            for (size_t p = 0; p < kid.size(); ++p) idx[fill[kid[p]]++] = (int)p;
            TH_HIP(hipMalloc(dptr, sizeof(int) * ptr.size()));
            TH_HIP(hipMalloc(didx, sizeof(int) * idx.size()));
            TH_HIP(hipMemcpy(*dptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
            TH_HIP(hipMemcpy(*didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
            return TH_OK;
        };
        if (int rc = csr(kid_im, &h->d_kim_ptr, &h->d_kim_idx)) return rc;
        if (int rc = csr(kid_re, &h->d_kre_ptr, &h->d_kre_idx)) return rc;
    }
    return TH_OK;
}

// Eigenpairs for the sides of th_chi0 at this q: set 0 (q = 0) always, set
// 1 for k + q when q != 0 (Jacobi in LDS for d <= 54, LAPACK's divide and
// conquer from rocSOLVER for d = 72, 144 and 288).  Skipped (the elimination
// kernels are used) when H is not Hermitian, d has no spectral kernel, or
// the sets would take more than 30 % of free HBM.
int th_ensure_eig(th_handle *h, const double *q)
{
    if (!h->spectral || !h->hermitian || !th_spectral_d(h->d) || h->n_kuniq <= 0) return TH_OK;
    const bool qz = q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0;
    const size_t sh = th_eig_split_halves(h->d);
    const size_t bytes = (size_t)h->n_kuniq * (h->d * (h->d * sizeof(double2) + sizeof(double)) + 2 * sh);
    // the handle's cap on its caches (both eigenpair sets count against it)
    if (h->cache_mb > 0 && (qz ? 1 : 2) * bytes > ((size_t)h->cache_mb << 20)) return TH_OK;
    for (int which = 0; which < (qz ? 1 : 2); ++which) {
        th_eig_set &es = h->eig[which];
        const double qq[3] = {which ? q[0] : 0.0, which ? q[1] : 0.0, which ? q[2] : 0.0};
        if (h->eig_valid[which] && es.q[0] == qq[0] && es.q[1] == qq[1] && es.q[2] == qq[2]) continue;
        h->eig_valid[which] = 0;
        if (es.V && es.cap < h->n_kuniq) {   // the k-point list grew: buffers of the new size
            TH_HIP(hipStreamSynchronize(h->stream));
            hipFree(es.V); hipFree(es.L); hipFree(es.S);
            es.V = nullptr; es.L = nullptr; es.S = nullptr;
            es.cap = 0;
        }
        if (!es.V) {
            size_t freeb = 0, totb = 0;
            if (hipMemGetInfo(&freeb, &totb) != hipSuccess || bytes > (size_t)(0.3 * (double)freeb)) return TH_OK;
            if (hipMalloc(&es.V, (size_t)h->n_kuniq * h->d * h->d * sizeof(double2)) != hipSuccess ||
                hipMalloc(&es.L, (size_t)h->n_kuniq * h->d * sizeof(double)) != hipSuccess ||
                (sh && hipMalloc(&es.S, (size_t)h->n_kuniq * sh * 2) != hipSuccess)) {
                hipFree(es.V); hipFree(es.L); hipFree(es.S);
                es.V = nullptr; es.L = nullptr; es.S = nullptr;
                (void)hipGetLastError();
                return TH_OK;
            }
            es.cap = h->n_kuniq;
        }
        int rc = th_launch_eig(h, which, qq);
        if (rc) return rc;
        es.q[0] = qq[0]; es.q[1] = qq[1]; es.q[2] = qq[2];
        h->eig_valid[which] = 1;
    }
    return TH_OK;
}

namespace {

long batch_pairs(th_handle *h, long total)
{
    const size_t per = (size_t)h->nA2 * 2 * h->Mpad * sizeof(double);
    const size_t budget = (size_t)h->batch_mb << 20;
    long b = (long)std::max<size_t>(TH_KC, budget / per);
    // a budget of at least one full wave of the fp16 spectral Green kernel's
    // workgroups is rounded to the nearest whole number of waves, so that no
    // launch ends with a part-empty wave
    const long qw = th_green_wave_pairs(h);
    if (qw > 0 && b >= qw) b = ((b + qw / 2) / qw) * qw;
    return std::min(b, std::max<long>(total, 1));
}

int fill_bcache(th_handle *h, double Ef, double etap, int prec, bool presplit, long geo, long bp, stage_timer &tm,
                bool *have);

// imaginary-axis B panels G(k, Ef + i(y+etap)) do not depend on omega:
// computed once per (Ef, etap, precision) and kept while they fit
int ensure_bcache(th_handle *h, double Ef, double etap, int prec, bool presplit, stage_timer &tm, bool *have)
{
    *have = false;
    // stored per batch of the imaginary contraction, rows in its summation order
    const long bp = batch_pairs(h, h->n_im), geo = h->pair_order ? bp : 0;
    if (h->d_bcache && h->bcache_prec == prec && h->bcache_key[0] == Ef && h->bcache_key[1] == etap &&
        h->bcache_bp == geo && h->bcache_presplit == (presplit ? 1 : 0)) {
        *have = true;
        return TH_OK;
    }
    h->bcache_prec = -1;
    const size_t need = (size_t)h->n_im * h->nA2 * 2 * h->Mpad * prec_size(prec);
    if (h->d_bcache && h->bcache_bytes != need) {
        hipFree(h->d_bcache);
        h->d_bcache = nullptr;
        h->bcache_bytes = 0;
    }
    if (h->d_bcache) return fill_bcache(h, Ef, etap, prec, presplit, geo, bp, tm, have);   // refill in place
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return TH_OK;
    // the cache may take most of HBM (288 GB on MI355X); otherwise recompute per
    // batch.  With a per-handle cap, the eigenpairs already held count against it.
    if (need > (size_t)(0.7 * (double)freeb)) return TH_OK;
    if (h->cache_mb > 0) {
        size_t held = 0;
        for (const th_eig_set &es : h->eig)
            if (es.V)
                held += (size_t)es.cap * (h->d * (h->d * sizeof(double2) + sizeof(double)) + 2 * th_eig_split_halves(h->d));
        if (need + held > ((size_t)h->cache_mb << 20)) return TH_OK;
    }
    if (hipMalloc(&h->d_bcache, need) != hipSuccess) {
        h->d_bcache = nullptr;
        (void)hipGetLastError();
        return TH_OK;
    }
    h->bcache_bytes = need;
    return fill_bcache(h, Ef, etap, prec, presplit, geo, bp, tm, have);
}

// the cache's panels computed into its allocation
int fill_bcache(th_handle *h, double Ef, double etap, int prec, bool presplit, long geo, long bp, stage_timer &tm,
                bool *have)
{
    double bb = 0.0;   // |G| <= 1/(y + eta') over the whole list
    if (presplit)
        for (double y : h->yim_host) bb = std::max(bb, 1.0 / (y + etap));
    for (long p0 = 0; p0 < h->n_im; p0 += bp) {
        th_green_args a{};
        a.mode = TH_G_IMAG_B;
        a.p0 = p0;
        a.np = std::min(bp, h->n_im - p0);
        a.Ef = Ef;
        a.etap = etap;
        a.prec = prec;
        a.sum_order = 1;
        if (presplit) {
            a.presplit = 1;
            a.pexp = th_pow2_scale(bb);
            h->bcache_pexp = a.pexp;
        }
        a.panel = (char *)h->d_bcache + (size_t)p0 * h->nA2 * 2 * h->Mpad * prec_size(prec);
        int id = tm.begin(0);
        int rc = th_launch_green(h, a);
        tm.end(id);
        if (rc) return rc;
    }
    h->bcache_prec = prec;
    h->bcache_key[0] = Ef;
    h->bcache_key[1] = etap;
    h->bcache_bp = geo;
    h->bcache_presplit = presplit ? 1 : 0;
    *have = true;
    return TH_OK;
}

// Host side of one omega of chi0: argument checks, real-axis nodes to the
// device, the omega-independent B cache, panel buffers.  Issues no kernel of
// the contraction itself.
int chi0_prepare(th_handle *h, const double *q, double omega, double Ef, double etap, int n_x, const double *x2,
                 const double *p2, int prec, stage_timer &tm, bool *real_axis, bool *have_b, bool *refine_on)
{
    *refine_on = false;
    if (!h->d_kim || h->n_im <= 0) {
        th_set_error("th_set_imag_points must be called before th_chi0");
        return TH_E_STATE;
    }
    *real_axis = std::fabs(omega) >= 1e-15 && n_x > 0;
    if (*real_axis && (!h->d_kre || h->nk_re <= 0)) {
        th_set_error("th_set_real_kmesh must be called before th_chi0 at omega != 0");
        return TH_E_STATE;
    }
    if (*real_axis && (long)n_x * h->nk_re >= (1L << 31)) {
        th_set_error("th_chi0: n_x * real k-points must stay below 2^31");
        return TH_E_ARG;
    }
    int rc;
    if (*real_axis && (rc = th_upload_nodes(h, n_x, x2, p2))) return rc;
    {
        const int id = tm.begin(0);
        rc = th_ensure_eig(h, q);
        tm.end(id);
        if (rc) return rc;
    }
    // split half-precision panels from the Green function kernels (also for
    // the cache) when both sides take the half-precision spectral kernel
    const bool presplit = prec == TH_PREC_MIXED && h->contract_f16 && h->hermitian && etap > 0.0 &&
                          std::isfinite(etap) && th_green_presplit_ok(h, q);
    if ((rc = ensure_bcache(h, Ef, etap, prec, presplit, tm, have_b))) return rc;
    if ((rc = th_refine_prepare(h, q, prec, refine_on))) return rc;
    long need = batch_pairs(h, h->n_im);
    if (*real_axis) need = std::max(need, batch_pairs(h, (long)n_x * h->nk_re));
    if ((rc = th_ensure_part(h, th_contract_splits(h, need, prec)))) return rc;
    return ensure_panels(h, need);
}

// Scales of the FP32 panels for the half-precision contraction from the
// resolvent bound |(z - H)^-1| <= 1/Im z of a Hermitian H: imaginary axis
// |w G_q| <= |w|/(y + eta) and |G| <= 1/(y + eta'); real axis
// |w (G_q - G_q^H)| <= 2 |p2| |w_k|/eta (the node weight p2 from device
// memory) and |G^H| <= 1/eta'.  False when a bound is not finite.
bool contract_scales(th_handle *h, double eta, double etap, int n_x, th_contract_scale *im, th_contract_scale *re)
{
    if (!h->hermitian || !(eta > 0.0) || !(etap > 0.0) || !std::isfinite(eta) || !std::isfinite(etap)) return false;
    if (h->im_bound_key[0] != eta || h->im_bound_key[1] != etap) {
        double ba = 0.0, bb = 0.0;
        for (size_t p = 0; p < h->yim_host.size(); ++p) {
            ba = std::max(ba, std::fabs(h->wim_host[p]) / (h->yim_host[p] + eta));
            bb = std::max(bb, 1.0 / (h->yim_host[p] + etap));
        }
        h->im_bound[0] = ba;
        h->im_bound[1] = bb;
        h->im_bound_key[0] = eta;
        h->im_bound_key[1] = etap;
    }
    for (double b : h->im_bound)
        if (!(b >= 0.0) || !std::isfinite(b)) return false;
    im->eA = th_pow2_scale(h->im_bound[0]);
    im->eB = th_pow2_scale(h->im_bound[1]);
    re->areal = 2.0 * h->wre_absmax / eta;
    if (!std::isfinite(re->areal)) return false;
    re->eA = 0;
    re->eB = th_pow2_scale(1.0 / etap);
    re->p2 = h->d_p2;
    re->n_x = n_x;
    return true;
}

// Kernel launches of one omega of chi0 into h->d_chi0 (everything on the
// handle's stream, no host synchronisation, no allocation: capturable).
// omega_p, when set, points to omega in device memory.
int chi0_launch(th_handle *h, const double *q, double omega, const double *omega_p, double Ef, double eta,
                double etap, int n_x, bool real_axis, bool have_b, int prec, stage_timer *tm, bool refine_on)
{
    const size_t nacc = (size_t)2 * h->nA2 * 2 * h->Mpad * h->Mpad;
    int rc = th_launch_zero(h, h->d_acc, nacc);
    if (rc) return rc;
    const size_t pstride = (size_t)h->nA2 * 2 * h->Mpad * prec_size(prec);
    // The inversions of a batch (VALU) and the contraction of the previous
    // batch (MFMA) are independent: with overlap on, the inversions go to
    // stream2 into one of two panel slots, and the contraction of slot s waits
    // for its inversions; the inversions into slot s wait for the contraction
    // that last read it.  The fork and join events make the two streams one
    // dependency graph, so the step can be captured as a hipGraph.
    // overlap 1 pipelines the elimination kernels only; 2 also the spectral
    // Green functions (their fp16 variant is issue-bound on the vector ALUs,
    // the contraction on the matrix cores and LDS)
    const bool spectral_on = h->spectral && h->eig_valid[0];
    h->last_spectral = spectral_on ? 1 : 0;
    const bool ov = h->overlap == 2 || (h->overlap == 1 && !spectral_on);
    hipStream_t gs = ov ? h->stream2 : h->stream;
    void *pA[2] = {h->d_pA, ov ? h->d_pA2 : h->d_pA}, *pB[2] = {h->d_pB, ov ? h->d_pB2 : h->d_pB};
    if (ov) {
        TH_HIP(hipEventRecord(h->ev_fork, h->stream));
        TH_HIP(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
    }
    long nb = 0;   // batches issued in this call
    th_contract_scale sc[2];
    const bool scaled = prec == TH_PREC_MIXED && contract_scales(h, eta, etap, n_x, &sc[0], &sc[1]);
    // both sides' Green function kernels write fp16 splits (with the
    // contraction's exponents); a cache of the other form is not used
    const bool presplit = scaled && h->contract_f16 && th_green_presplit_ok(h, q);
    sc[0].presplit = sc[1].presplit = presplit;
    h->last_presplit = presplit ? 1 : 0;
    if (have_b && h->bcache_presplit != (presplit ? 1 : 0)) have_b = false;
    auto begin = [&](int st, hipStream_t s) { return tm ? tm->begin(st, s) : -1; };
    auto end = [&](int id) { if (tm) tm->end(id); };
    auto batch = [&](int mode_a, int mode_b, long p0, long np, int axis) -> int {
        const int slot = ov ? (int)(nb & 1) : 0;
        if (ov && nb >= 2) TH_HIP(hipStreamWaitEvent(h->stream2, h->ev_contract[slot], 0));
        th_green_args a{};
        a.p0 = p0; a.np = np; a.q[0] = q[0]; a.q[1] = q[1]; a.q[2] = q[2];
        a.omega = omega; a.omega_p = omega_p; a.Ef = Ef; a.eta = eta; a.etap = etap; a.prec = prec;
        a.stream = gs;
        a.sum_order = 1;   // panel rows in the contraction's summation order
        if (axis == 1) {   // real axis in k-major pair order (runs of n_x pairs per k-point)
            a.kmajor_re = h->pair_order;
            a.nx = n_x;
        }
        a.mode = mode_a;
        a.panel = pA[slot];
        if (presplit) {
            a.presplit = 1;
            a.pexp = sc[axis].eA;
            if (axis == 1) {   // exponent from the node weights in device memory
                a.pareal = sc[1].areal;
                a.pnx = n_x;
            }
        }
        int id = begin(0, gs);
        int r = th_launch_green(h, a);
        if (r) return r;
        const void *b = nullptr;
        if (mode_b < 0) {
            b = (const char *)h->d_bcache + (size_t)p0 * pstride;
        } else {
            a.mode = mode_b;
            a.panel = pB[slot];
            a.pexp = sc[axis].eB;
            a.pareal = 0.0;
            if ((r = th_launch_green(h, a))) return r;
            b = pB[slot];
        }
        end(id);
        if (ov) {
            TH_HIP(hipEventRecord(h->ev_green[slot], h->stream2));
            TH_HIP(hipStreamWaitEvent(h->stream, h->ev_green[slot], 0));
        }
        id = begin(1, h->stream);
        if ((r = th_launch_contract(h, pA[slot], b, np, prec, axis, scaled ? &sc[axis] : nullptr))) return r;
        end(id);
        if (ov) TH_HIP(hipEventRecord(h->ev_contract[slot], h->stream));
        ++nb;
        return TH_OK;
    };
    // imaginary axis
    const long bp = batch_pairs(h, h->n_im);
    for (long p0 = 0; p0 < h->n_im; p0 += bp)
        if ((rc = batch(TH_G_IMAG_A, have_b ? -1 : TH_G_IMAG_B, p0, std::min(bp, h->n_im - p0), 0))) return rc;
    // real axis
    if (real_axis) {
        const long n_re = (long)n_x * h->nk_re;
        const long br = batch_pairs(h, n_re);
        for (long p0 = 0; p0 < n_re; p0 += br)
            if ((rc = batch(TH_G_REAL_A, TH_G_REAL_B, p0, std::min(br, n_re - p0), 1))) return rc;
    }
    // join: every inversion of this step is done before the finalize (the
    // contractions on 'stream' already waited for them)
    if (ov && nb > 0) TH_HIP(hipStreamWaitEvent(h->stream, h->ev_green[(nb - 1) & 1], 0));
    int id = begin(2, h->stream);
    rc = th_launch_finalize(h, h->d_chi0);
    // the block chi0[S,S] that the Dyson step amplifies near a pole, in FP64
    if (!rc && refine_on) rc = th_refine_launch(h, q, omega, omega_p, Ef, eta, etap, n_x, real_axis, h->d_chi0);
    h->last_refined = refine_on ? 1 : 0;
    end(id);
    return rc;
}

// chi0 of one omega into h->d_chi0 (device)
int chi0_device(th_handle *h, const double *q, double omega, double Ef, double eta, double etap, int n_x,
                const double *x2, const double *p2, int prec, stage_timer &tm)
{
    bool real_axis = false, have_b = false, refine_on = false;
    int rc = chi0_prepare(h, q, omega, Ef, etap, n_x, x2, p2, prec, tm, &real_axis, &have_b, &refine_on);
    if (rc) return rc;
    return chi0_launch(h, q, omega, nullptr, Ef, eta, etap, n_x, real_axis, have_b, prec, &tm, refine_on);
}

bool bad(const void *p) { return p == nullptr; }

}  // namespace

extern "C" {

const char *th_last_error(void) { return g_err.c_str(); }

int th_limits(int *max_nOrb, int *max_d, int *max_nA)
{
    if (max_nOrb) *max_nOrb = TH_MAX_NORB;
    if (max_d) *max_d = TH_MAX_D;
    if (max_nA) *max_nA = TH_MAX_NA;
    return TH_OK;
}

int th_create(int nA, int nOrb, int nhop, const double *R, const int *jb, const int *ib, const double *t,
              const double *h0, th_handle **out)
{
    if (!out) { th_set_error("th_create: out is NULL"); return TH_E_ARG; }
    *out = nullptr;
    if (nA < 1 || nA > TH_MAX_NA || nOrb < 1 || nOrb > TH_MAX_NORB || nhop < 0 || bad(h0) ||
        (nhop > 0 && (bad(R) || bad(jb) || bad(ib) || bad(t)))) {
        th_set_error("th_create: bad sizes or NULL arrays");
        return TH_E_ARG;
    }
    for (int n = 0; n < nhop; ++n)
        if (jb[n] < 0 || jb[n] >= nA || ib[n] < 0 || ib[n] >= nA) {
            th_set_error("th_create: hop endpoint outside 0..nA-1 at entry " + std::to_string(n));
            return TH_E_ARG;
        }
    const int d = 2 * nOrb * nA;
    if (d > TH_MAX_D_LDS && d != 72 && d != 144 && d != 288) {
        th_set_error("th_create: d = 2*nOrb*nA = " + std::to_string(d) + " exceeds " + std::to_string(TH_MAX_D_LDS));
        return TH_E_UNSUPPORTED;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        th_set_error("th_create: no HIP device");
        return TH_E_NODEVICE;
    }
    th_handle *h = new th_handle();
    TH_HIP(hipGetDevice(&h->device));
    // the allocation-free Dyson entry point's scratch on this device
    if (int rc = th_noalloc_init(h->device)) {
        delete h;
        return rc;
    }
    hipDeviceProp_t prop;
    TH_HIP(hipGetDeviceProperties(&prop, h->device));
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
        th_set_error(std::string("th_create: device is ") + prop.gcnArchName + ", this library is built for gfx950");
        delete h;
        return TH_E_NODEVICE;
    }
    TH_HIP(hipStreamCreate(&h->stream));
    TH_HIP(hipStreamCreate(&h->stream2));
    TH_HIP(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    for (int s = 0; s < 2; ++s) {
        TH_HIP(hipEventCreateWithFlags(&h->ev_green[s], hipEventDisableTiming));
        TH_HIP(hipEventCreateWithFlags(&h->ev_contract[s], hipEventDisableTiming));
    }
    h->nA = nA; h->nOrb = nOrb; h->bs = 2 * nOrb; h->d = d; h->D = 4 * nOrb * nOrb * nA; h->nhop = nhop;
    h->M = h->bs * h->bs; h->Mpad = th_mpad(h->M); h->nA2 = nA * nA;
    const int ntile = h->Mpad / TH_TILE;
    h->splits = std::max(1, (2 * prop.multiProcessorCount) / (ntile * ntile * h->nA2));
    int *st = nullptr;
    TH_HIP(hipMalloc(&st, sizeof(int)));
    {
        std::lock_guard<std::mutex> lk(g_status_mu);
        g_status.push_back({h, st});
    }
    // hops grouped by atom pair (jb, ib) so an entry of H(k) visits only its own hops
    std::vector<int> order(nhop), hoff(nA * nA + 1, 0);
    for (int n = 0; n < nhop; ++n) {
        order[n] = n;
        hoff[jb[n] * nA + ib[n] + 1] += 1;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return jb[a] * nA + ib[a] < jb[b] * nA + ib[b]; });
    for (int k = 0; k < nA * nA; ++k) hoff[k + 1] += hoff[k];
    h->hop_order = order;
    TH_HIP(hipMalloc(&h->d_hoff, sizeof(int) * (nA * nA + 1)));
    TH_HIP(hipMemcpy(h->d_hoff, hoff.data(), sizeof(int) * (nA * nA + 1), hipMemcpyHostToDevice));
    if (nhop > 0) {
        const size_t nO2 = (size_t)nOrb * nOrb;
        std::vector<double> Rs(3 * (size_t)nhop), ts(2 * nO2 * nhop);
        std::vector<int> jbs(nhop), ibs(nhop);
        for (int n = 0; n < nhop; ++n) {
            const int o = order[n];
            for (int x = 0; x < 3; ++x) Rs[3 * n + x] = R[3 * o + x];
            jbs[n] = jb[o];
            ibs[n] = ib[o];
            std::memcpy(&ts[2 * nO2 * n], &t[2 * nO2 * o], sizeof(double) * 2 * nO2);
        }
        TH_HIP(hipMalloc(&h->d_R, sizeof(double) * 3 * nhop));
        TH_HIP(hipMalloc(&h->d_jb, sizeof(int) * nhop));
        TH_HIP(hipMalloc(&h->d_ib, sizeof(int) * nhop));
        TH_HIP(hipMalloc(&h->d_t, sizeof(double2) * (size_t)nhop * nOrb * nOrb));
        TH_HIP(hipMemcpy(h->d_R, Rs.data(), sizeof(double) * 3 * nhop, hipMemcpyHostToDevice));
        TH_HIP(hipMemcpy(h->d_jb, jbs.data(), sizeof(int) * nhop, hipMemcpyHostToDevice));
        TH_HIP(hipMemcpy(h->d_ib, ibs.data(), sizeof(int) * nhop, hipMemcpyHostToDevice));
        TH_HIP(hipMemcpy(h->d_t, ts.data(), sizeof(double2) * (size_t)nhop * nOrb * nOrb, hipMemcpyHostToDevice));
    }
    // H(k) is Hermitian for every k when h0 is and the hops come in pairs
    // (R, jb, ib, t) and (-R, ib, jb, t^H)
    {
        const double2 *H0 = reinterpret_cast<const double2 *>(h0);
        double scale = 0.0;
        for (size_t e = 0; e < (size_t)d * d; ++e) scale = std::max(scale, std::fabs(H0[e].x) + std::fabs(H0[e].y));
        const double tol = 1e-12 * std::max(scale, 1.0);
        bool herm = true;
        for (int c = 0; c < d && herm; ++c)
            for (int r = 0; r < d; ++r) {
                const double2 a = H0[r + (size_t)d * c], b = H0[c + (size_t)d * r];
                if (std::fabs(a.x - b.x) > tol || std::fabs(a.y + b.y) > tol) { herm = false; break; }
            }
        const size_t nO2 = (size_t)nOrb * nOrb;
        const double2 *T = reinterpret_cast<const double2 *>(t);
        for (int n = 0; n < nhop && herm; ++n) {
            bool found = false;
            for (int m = 0; m < nhop && !found; ++m) {
                if (jb[m] != ib[n] || ib[m] != jb[n]) continue;
                if (std::fabs(R[3 * m] + R[3 * n]) > 1e-9 || std::fabs(R[3 * m + 1] + R[3 * n + 1]) > 1e-9 ||
                    std::fabs(R[3 * m + 2] + R[3 * n + 2]) > 1e-9)
                    continue;
                bool same = true;
                for (int nu = 0; nu < nOrb && same; ++nu)
                    for (int mu = 0; mu < nOrb; ++mu) {
                        const double2 x = T[n * nO2 + nu + (size_t)nOrb * mu], y = T[m * nO2 + mu + (size_t)nOrb * nu];
                        if (std::fabs(x.x - y.x) > tol || std::fabs(x.y + y.y) > tol) { same = false; break; }
                    }
                found = same;
            }
            herm = found;
        }
        h->hermitian = herm ? 1 : 0;
    }
    TH_HIP(hipMalloc(&h->d_h0, sizeof(double2) * (size_t)d * d));
    TH_HIP(hipMemcpy(h->d_h0, h0, sizeof(double2) * (size_t)d * d, hipMemcpyHostToDevice));
    const size_t MM = (size_t)h->Mpad * h->Mpad;
    TH_HIP(hipMalloc(&h->d_part, sizeof(double) * (size_t)h->splits * h->nA2 * 2 * MM));
    h->part_cap = h->splits;
    TH_HIP(hipMalloc(&h->d_acc, sizeof(double) * (size_t)2 * h->nA2 * 2 * MM));
    TH_HIP(hipMalloc(&h->d_chi0, sizeof(double2) * (size_t)h->D * h->D));
    TH_HIP(hipMalloc(&h->d_chi, sizeof(double2) * (size_t)h->D * h->D));
    *out = h;
    return TH_OK;
}

int th_set_imag_points(th_handle *h, long n, const double *k, const double *y, const double *w)
{
    if (!h || n < 1 || bad(k) || bad(y) || bad(w)) { th_set_error("th_set_imag_points: bad arguments"); return TH_E_ARG; }
    if (n >= (1L << 31)) { th_set_error("th_set_imag_points: at most 2^31 - 1 pairs"); return TH_E_ARG; }
    hipFree(h->d_kim); hipFree(h->d_yim); hipFree(h->d_wim); hipFree(h->d_bcache);
    h->d_kim = h->d_yim = h->d_wim = nullptr;
    h->d_bcache = nullptr;
    h->bcache_bytes = 0;
    h->bcache_prec = -1;
    // unique k-points of the list
    std::unordered_map<kkey, int, kkey_hash> ids;
    h->kim_uniq.clear();
    std::vector<int> kid(n);
    for (long p = 0; p < n; ++p) {
        auto key = make_key(k + 3 * p);
        auto it = ids.find(key);
        if (it == ids.end()) {
            const int id = (int)(h->kim_uniq.size() / 3);
            ids.emplace(key, id);
            h->kim_uniq.insert(h->kim_uniq.end(), k + 3 * p, k + 3 * p + 3);
            kid[p] = id;
        } else {
            kid[p] = it->second;
        }
    }
    // pair order on the device: as given, or in runs of TH_RUN pairs of one
    // k-point: the r-th pair of a k-point (in the given order) goes to block
    // r / TH_RUN, and the list is sorted by (block, k-point, r).  Each run
    // shares one eigen-decomposition (the spectral kernel loads V once per
    // run) and consecutive runs belong to different k-points, so that a
    // batch holds few pairs of any one k-point.  chi0 is a sum over the
    // pairs; the order changes only the rounding of that sum.
    std::vector<long> order(n);
    for (long p = 0; p < n; ++p) order[p] = p;
    if (h->pair_order) {
        const size_t nu = h->kim_uniq.size() / 3;
        std::vector<long> seen(nu, 0), blk(n);
        long nblk = 0;
        for (long p = 0; p < n; ++p) {
            blk[p] = seen[kid[p]]++ / TH_RUN;
            nblk = std::max(nblk, blk[p] + 1);
        }
        // counting sort by k-point, then (stable) by block
        std::vector<long> start(nu + 1, 0), byk(n);
        for (long p = 0; p < n; ++p) ++start[kid[p] + 1];
        for (size_t u = 1; u < start.size(); ++u) start[u] += start[u - 1];
        for (long p = 0; p < n; ++p) byk[start[kid[p]]++] = p;
        std::vector<long> bstart(nblk + 1, 0);
        for (long p = 0; p < n; ++p) ++bstart[blk[p] + 1];
        for (long b = 1; b <= nblk; ++b) bstart[b] += bstart[b - 1];
        for (long i = 0; i < n; ++i) order[bstart[blk[byk[i]]]++] = byk[i];
    }
    std::vector<double> kk(3 * n), yy(n), ww(n);
    h->kid_im_local.assign(n, 0);
    for (long s = 0; s < n; ++s) {
        const long p = order[s];
        kk[3 * s] = k[3 * p]; kk[3 * s + 1] = k[3 * p + 1]; kk[3 * s + 2] = k[3 * p + 2];
        yy[s] = y[p];
        ww[s] = w[p];
        h->kid_im_local[s] = kid[p];
    }
    TH_HIP(hipMalloc(&h->d_kim, sizeof(double) * 3 * n));
    TH_HIP(hipMalloc(&h->d_yim, sizeof(double) * n));
    TH_HIP(hipMalloc(&h->d_wim, sizeof(double) * n));
    TH_HIP(hipMemcpy(h->d_kim, kk.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(h->d_yim, yy.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(h->d_wim, ww.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    h->n_im = n;
    h->yim_host = std::move(yy);
    h->wim_host = std::move(ww);
    h->im_bound_key[0] = h->im_bound_key[1] = -1.0;
    return rebuild_kunion(h);
}

int th_set_real_kmesh(th_handle *h, long nk, const double *k, const double *w)
{
    if (!h || nk < 1 || bad(k) || bad(w)) { th_set_error("th_set_real_kmesh: bad arguments"); return TH_E_ARG; }
    hipFree(h->d_kre); hipFree(h->d_wre);
    h->d_kre = h->d_wre = nullptr;
    TH_HIP(hipMalloc(&h->d_kre, sizeof(double) * 3 * nk));
    TH_HIP(hipMalloc(&h->d_wre, sizeof(double) * nk));
    TH_HIP(hipMemcpy(h->d_kre, k, sizeof(double) * 3 * nk, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(h->d_wre, w, sizeof(double) * nk, hipMemcpyHostToDevice));
    h->nk_re = nk;
    h->kre_host.assign(k, k + 3 * nk);
    h->wre_absmax = 0.0;
    for (long i = 0; i < nk; ++i) h->wre_absmax = std::max(h->wre_absmax, std::fabs(w[i]));
    return rebuild_kunion(h);
}

int th_chi0(th_handle *h, const double *q, double omega, double Ef, double eta, double etap, int n_x,
            const double *x2, const double *p2, int precision, double *chi0_out)
{
    if (!h || bad(q) || bad(chi0_out) || (precision != TH_PREC_FP64 && precision != TH_PREC_MIXED) ||
        (std::fabs(omega) >= 1e-15 && n_x > 0 && (bad(x2) || bad(p2))) || n_x < 0) {
        th_set_error("th_chi0: bad arguments");
        return TH_E_ARG;
    }
    stage_timer tm(h);
    int id = tm.begin(4);
    for (auto &c : h->n_launch) c = 0;
    int rc = th_reset_status(h);
    if (!rc) rc = chi0_device(h, q, omega, Ef, eta, etap, n_x, x2, p2, precision, tm);
    tm.end(id);
    if (rc) return rc;
    TH_HIP(hipMemcpyAsync(chi0_out, h->d_chi0, sizeof(double2) * (size_t)h->D * h->D, hipMemcpyDeviceToHost,
                          h->stream));
    rc = th_check_status(h);
    tm.collect();
    return rc;
}

int th_dyson(th_handle *h, const double *chi0, const double *U, double *chi_out, double *cond_est)
{
    if (!h || bad(chi0) || bad(U) || bad(chi_out)) { th_set_error("th_dyson: bad arguments"); return TH_E_ARG; }
    const size_t n = (size_t)h->D * h->D;
    int rc = th_reset_status(h);
    if (rc) return rc;
    TH_HIP(hipMemcpyAsync(h->d_chi0, chi0, sizeof(double2) * n, hipMemcpyHostToDevice, h->stream));
    stage_timer tm(h);
    int id = tm.begin(3);
    rc = th_dyson_device(h, h->d_chi0, reinterpret_cast<const double2 *>(U), h->d_chi, cond_est);
    tm.end(id);
    if (rc) return rc;
    TH_HIP(hipMemcpyAsync(chi_out, h->d_chi, sizeof(double2) * n, hipMemcpyDeviceToHost, h->stream));
    rc = th_check_status(h);
    tm.collect();
    return rc;
}

int th_dyson_noalloc(th_handle *h, const double *chi0_dev, const double *U, double *chi_dev, double *cond_est)
{
    if (!h || bad(chi0_dev) || bad(U) || bad(chi_dev)) {
        th_set_error("th_dyson_noalloc: bad arguments");
        return TH_E_ARG;
    }
    return th_dyson_noalloc_impl(h, reinterpret_cast<const double2 *>(chi0_dev), reinterpret_cast<const double2 *>(U),
                                 reinterpret_cast<double2 *>(chi_dev), cond_est);
}

int th_set_interaction(th_handle *h, const double *U)
{
    if (!h) { th_set_error("th_set_interaction: bad arguments"); return TH_E_ARG; }
    return th_refine_set_support(h, reinterpret_cast<const double2 *>(U));
}

namespace {

// One omega-step of the sweep captured as a hipGraph: chi0 launches with
// omega read from device memory, the Dyson kernels and both traces.  All
// host work (nodes, caches, buffers, U support) is done before capture, so
// the graph only references buffers that stay in place for the whole sweep.
int capture_step(th_handle *h, const double *q, double Ef, double eta, double etap, int n_x, bool real_axis,
                 bool have_b, bool refine_on, int prec, int r, double2 *d_s, hipGraphExec_t *out)
{
    const int n4 = 4 * h->nA;
    hipGraph_t graph = nullptr;
    TH_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = chi0_launch(h, q, 0.0, h->d_omega, Ef, eta, etap, n_x, real_axis, have_b, prec, nullptr, refine_on);
    if (!rc) rc = th_dyson_launch(h, r, h->d_chi0, h->d_chi, h->d_norms);
    if (!rc) rc = th_launch_trace(h, h->d_chi0, d_s);
    if (!rc) rc = th_launch_trace(h, h->d_chi, d_s + n4 * n4);
    hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) {
        th_set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return TH_E_HIP;
    }
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
        th_set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return TH_E_HIP;
    }
    return TH_OK;
}

}  // namespace

int th_chi_sweep(th_handle *h, const double *q, int n_omega, const double *omega, double Ef, double eta,
                 double etap, int parts3, int n3gl, const double *U, int precision, double *schi0_out,
                 double *schi_out, double *cond_out, double *chi_full_out)
{
    if (!h || bad(q) || n_omega < 1 || bad(omega) || parts3 < 1 || n3gl < 1 || bad(U) ||
        (precision != TH_PREC_FP64 && precision != TH_PREC_MIXED)) {
        th_set_error("th_chi_sweep: bad arguments");
        return TH_E_ARG;
    }
    const int n4 = 4 * h->nA, n_x = parts3 * n3gl;
    std::vector<double> x2(n_x), p2(n_x), om(omega, omega + n_omega);
    double2 *d_s = nullptr;
    TH_HIP(hipMalloc(&d_s, sizeof(double2) * 2 * n4 * n4));
    if (h->use_graph && (!h->d_omega || !h->d_norms)) {
        if (!h->d_omega) TH_HIP(hipMalloc(&h->d_omega, sizeof(double)));
        if (!h->d_norms) TH_HIP(hipMalloc(&h->d_norms, 2 * sizeof(double)));
    }
    hipGraphExec_t ge[2] = {nullptr, nullptr};
    // device buffers a captured step reads or writes; a replay is valid only
    // while every one of them is still where it was at capture
    auto buffers = [h]() {
        return std::vector<const void *>{h->d_pA, h->d_pB, h->d_pA2, h->d_pB2, h->d_part, h->d_bcache, h->d_x2,
                                         h->d_p2, h->eig[0].V, h->eig[0].L, h->eig[1].V, h->eig[1].L, h->eig[0].S,
                                         h->eig[1].S,
                                         h->d_ref_int, h->d_ref_part};
    };
    std::vector<const void *> ge_buf[2];
    stage_timer tm(h);
    int idall = tm.begin(4);
    for (auto &c : h->n_launch) c = 0;
    int r = 0;
    // a sweep diagonalises H(k) afresh (once, at its first frequency)
    h->eig_valid[0] = h->eig_valid[1] = 0;
    int rc = th_reset_status(h);
    // the sweep's U is the one the mixed chi0 refines its support block for
    if (!rc && (precision == TH_PREC_MIXED || h->refine == 2))
        rc = th_refine_set_support(h, reinterpret_cast<const double2 *>(U));
    if (!rc && h->use_graph) rc = th_dyson_prepare(h, reinterpret_cast<const double2 *>(U), &r);
    if (!rc && h->use_graph) {
        // panels and partial sums sized for the largest batch of any frequency
        // of the list before the first capture, so that no later frequency
        // reallocates a buffer a captured step refers to
        bool any_real = false;
        for (double e : om) any_real = any_real || std::fabs(e) >= 1e-15;
        long need = batch_pairs(h, h->n_im);
        if (any_real && h->nk_re > 0) need = std::max(need, batch_pairs(h, (long)n_x * h->nk_re));
        rc = th_ensure_part(h, th_contract_splits(h, need, precision));
        if (!rc) rc = th_ensure_panels(h, need);
    }
    for (int w = 0; w < n_omega && !rc; ++w) {
        const double e = om[w];
        if (std::fabs(e) >= 1e-15)
            for (int k = 1; k <= parts3; ++k) {
                double e1 = Ef - ((parts3 - (k - 1)) * e / parts3);
                double e2 = Ef - ((parts3 - k) * e / parts3);
                gauleg(e1, e2, n3gl, x2.data() + (k - 1) * n3gl, p2.data() + (k - 1) * n3gl);
            }
        double cond = 0.0;
        if (h->use_graph) {
            bool real_axis = false, have_b = false, refine_on = false;
            rc = chi0_prepare(h, q, e, Ef, etap, n_x, x2.data(), p2.data(), precision, tm, &real_axis, &have_b,
                              &refine_on);
            if (rc) break;
            if (hipMemcpyAsync(h->d_omega, &om[w], sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
                th_set_error("th_chi_sweep: omega upload failed");
                rc = TH_E_HIP;
                break;
            }
            const int kind = real_axis ? 1 : 0;
            if (ge[kind] && ge_buf[kind] != buffers()) {   // a buffer moved since the capture
                hipGraphExecDestroy(ge[kind]);
                ge[kind] = nullptr;
            }
            if (!ge[kind]) {
                if ((rc = capture_step(h, q, Ef, eta, etap, n_x, real_axis, have_b, refine_on, precision, r, d_s,
                                       &ge[kind])))
                    break;
                ge_buf[kind] = buffers();
            }
            if (hipGraphLaunch(ge[kind], h->stream) != hipSuccess) {
                th_set_error("th_chi_sweep: hipGraphLaunch failed");
                rc = TH_E_HIP;
                break;
            }
            h->n_launch[0] += 1;
            if (cond_out) {
                double nv[2];
                if (hipMemcpyAsync(nv, h->d_norms, sizeof(nv), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                    hipStreamSynchronize(h->stream) != hipSuccess) {
                    th_set_error("th_chi_sweep: norm download failed");
                    rc = TH_E_HIP;
                    break;
                }
                cond = nv[0] * nv[1];
            }
        } else {
            rc = chi0_device(h, q, e, Ef, eta, etap, n_x, x2.data(), p2.data(), precision, tm);
            if (rc) break;
            int id = tm.begin(3);
            rc = th_dyson_device(h, h->d_chi0, reinterpret_cast<const double2 *>(U), h->d_chi,
                                 cond_out ? &cond : nullptr);
            if (!rc) rc = th_launch_trace(h, h->d_chi0, d_s);
            if (!rc) rc = th_launch_trace(h, h->d_chi, d_s + n4 * n4);
            tm.end(id);
            if (rc) break;
        }
        if (cond_out) cond_out[w] = cond;
        hipError_t ce = hipSuccess;
        if (schi0_out)
            ce = hipMemcpyAsync(schi0_out + (size_t)w * 2 * n4 * n4, d_s, sizeof(double2) * n4 * n4,
                                hipMemcpyDeviceToHost, h->stream);
        if (schi_out && ce == hipSuccess)
            ce = hipMemcpyAsync(schi_out + (size_t)w * 2 * n4 * n4, d_s + n4 * n4, sizeof(double2) * n4 * n4,
                                hipMemcpyDeviceToHost, h->stream);
        if (ce == hipSuccess) ce = hipStreamSynchronize(h->stream);
        if (ce != hipSuccess) {
            th_set_error(std::string("th_chi_sweep: ") + hipGetErrorString(ce));
            rc = TH_E_HIP;
        }
    }
    tm.end(idall);
    if (!rc && chi_full_out &&
        hipMemcpyAsync(chi_full_out, h->d_chi, sizeof(double2) * (size_t)h->D * h->D, hipMemcpyDeviceToHost,
                       h->stream) != hipSuccess) {
        th_set_error("th_chi_sweep: chi download failed");
        rc = TH_E_HIP;
    }
    if (!rc) rc = th_check_status(h);
    hipStreamSynchronize(h->stream);
    tm.collect();
    for (auto &g : ge)
        if (g) hipGraphExecDestroy(g);
    hipFree(d_s);
    return rc;
}

int th_last_timing(th_handle *h, double *t, long *n)
{
    if (!h) return TH_E_ARG;
    if (t)
        for (int s = 0; s < 5; ++s) t[s] = h->t_ms[s];
    if (n)
        for (int s = 0; s < 4; ++s) n[s] = h->n_launch[s];
    return TH_OK;
}


int th_tune(th_handle *h, const char *key, long value)
{
    if (!h || !key) { th_set_error("th_tune: bad arguments"); return TH_E_ARG; }
    const std::string k(key);
    if (k == "green_variant") {
        if (value < 0 || value > 3) { th_set_error("th_tune: green_variant is 0..3"); return TH_E_ARG; }
        h->kernel_variant = (int)value;
    } else if (k == "splits") {
        if (value < 1 || value > 4096) { th_set_error("th_tune: splits in 1..4096"); return TH_E_ARG; }
        h->splits = (int)value;
        if (int rc = th_ensure_part(h, (int)value)) return rc;
    } else if (k == "spectral") {
        if (value < 0 || value > 1) { th_set_error("th_tune: spectral is 0 or 1"); return TH_E_ARG; }
        h->spectral = (int)value;
    } else if (k == "spectral_chunk") {
        if (value < 0 || value > 1) { th_set_error("th_tune: spectral_chunk is 0 or 1"); return TH_E_ARG; }
        h->spectral_chunk = (int)value;
    } else if (k == "presplit") {
        if (value < 0 || value > 1) { th_set_error("th_tune: presplit is 0 or 1"); return TH_E_ARG; }
        h->presplit = (int)value;
    } else if (k == "green_f16") {
        if (value < 0 || value > 1) { th_set_error("th_tune: green_f16 is 0 or 1"); return TH_E_ARG; }
        h->green_f16 = (int)value;
    } else if (k == "contract_f16") {
        if (value < 0 || value > 1) { th_set_error("th_tune: contract_f16 is 0 or 1"); return TH_E_ARG; }
        h->contract_f16 = (int)value;
    } else if (k == "pair_order") {
        if (value < 0 || value > 1) { th_set_error("th_tune: pair_order is 0 or 1"); return TH_E_ARG; }
        h->pair_order = (int)value;   // applies from the next th_set_imag_points
    } else if (k == "overlap") {
        if (value < 0 || value > 2) { th_set_error("th_tune: overlap is 0, 1 or 2"); return TH_E_ARG; }
        h->overlap = (int)value;
    } else if (k == "current_general") {
        if (value < 0 || value > 1) { th_set_error("th_tune: current_general is 0 or 1"); return TH_E_ARG; }
        h->current_general = (int)value;
    } else if (k == "contract_tiles") {
        if (value < 0 || value > 1) { th_set_error("th_tune: contract_tiles is 0 or 1"); return TH_E_ARG; }
        h->contract_tiles = (int)value;
    } else if (k == "contract_3m") {
        if (value < 0 || value > 1) { th_set_error("th_tune: contract_3m is 0 or 1"); return TH_E_ARG; }
        h->contract_3m = (int)value;
    } else if (k == "graph") {
        if (value < 0 || value > 1) { th_set_error("th_tune: graph is 0 or 1"); return TH_E_ARG; }
        h->use_graph = (int)value;
    } else if (k == "drop_caches") {
        // forgets the omega-independent panel cache and the eigenpairs: the
        // next call computes them again (into the allocations it keeps)
        if (value != 1) { th_set_error("th_tune: drop_caches takes 1"); return TH_E_ARG; }
        TH_HIP(hipStreamSynchronize(h->stream));
        h->bcache_prec = -1;
        h->eig_valid[0] = h->eig_valid[1] = 0;
    } else if (k == "cache_mb") {
        if (value < 0) { th_set_error("th_tune: cache_mb >= 0"); return TH_E_ARG; }
        h->cache_mb = value;
    } else if (k == "refine") {
        if (value < 0 || value > 2) { th_set_error("th_tune: refine is 0, 1 or 2"); return TH_E_ARG; }
        h->refine = (int)value;
    } else if (k == "batch_mb") {
        if (value < 16 || value > 65536) { th_set_error("th_tune: batch_mb in 16..65536"); return TH_E_ARG; }
        h->batch_mb = (int)value;
    } else if (k == "adjoint_bcache") {
        if (value < 0 || value > 1) { th_set_error("th_tune: adjoint_bcache is 0 or 1"); return TH_E_ARG; }
        h->adj_bcache = (int)value;
    } else {
        th_set_error("th_tune: unknown key " + k);
        return TH_E_ARG;
    }
    return TH_OK;
}

int th_get_tuning(th_handle *h, long *green_variant, long *splits, long *batch_mb)
{
    if (!h) return TH_E_ARG;
    if (green_variant) *green_variant = h->kernel_variant;
    if (splits) *splits = h->splits;
    if (batch_mb) *batch_mb = h->batch_mb;
    return TH_OK;
}

int th_destroy(th_handle *h)
{
    if (!h) return TH_OK;
    hipStreamSynchronize(h->stream);
    th_eig_release(h);
    void *bufs[] = {h->d_hoff, h->d_R, h->d_jb, h->d_ib, h->d_t, h->d_h0, h->d_kim, h->d_yim, h->d_wim, h->d_kre,
                    h->d_wre, h->d_x2, h->d_p2, h->d_pA, h->d_pB, h->d_pA2, h->d_pB2, h->d_part, h->d_acc, h->d_chi0, h->d_chi,
                    h->d_bcache, h->d_sup, h->d_U, h->d_W, h->d_V, h->d_Z, h->d_X, h->d_adj, h->d_omega, h->d_norms, h->d_blk, h->d_kuniq, h->d_kid_im, h->d_kid_re,
                    h->eig[0].V, h->eig[0].L, h->eig[1].V, h->eig[1].L, h->eig[0].S, h->eig[1].S, h->d_kim_ptr, h->d_kim_idx,
                    h->d_kre_ptr, h->d_kre_idx, h->d_ref_int, h->d_ref_part, h->d_pamax, h->d_pays, h->d_wk};
    for (void *b : bufs)
        if (b) hipFree(b);
    {
        std::lock_guard<std::mutex> lk(g_status_mu);
        for (size_t i = 0; i < g_status.size(); ++i)
            if (g_status[i].first == h) {
                hipFree(g_status[i].second);
                g_status.erase(g_status.begin() + i);
                break;
            }
    }
    hipStreamSynchronize(h->stream2);
    hipStreamDestroy(h->stream2);
    hipEventDestroy(h->ev_fork);
    for (int s = 0; s < 2; ++s) {
        hipEventDestroy(h->ev_green[s]);
        hipEventDestroy(h->ev_contract[s]);
    }
    hipStreamDestroy(h->stream);
    delete h;
    return TH_OK;
}


/* ---- kernel-level test hooks (used by the GPU unit tests) ---- */

int thx_green_panel(th_handle *h, int mode, long p0, long np, const double *q, double omega, double Ef,
                    double eta, double etap, int prec, int n_x, const double *x2, const double *p2, double *out)
{
    if (!h || np < 1 || bad(q) || bad(out)) { th_set_error("thx_green_panel: bad arguments"); return TH_E_ARG; }
    long total = (mode <= TH_G_IMAG_B) ? h->n_im : (long)n_x * h->nk_re;
    if (p0 < 0 || p0 + np > total) { th_set_error("thx_green_panel: pair range"); return TH_E_ARG; }
    int rc = th_reset_status(h);
    if (rc) return rc;
    if (mode >= TH_G_REAL_A && (rc = th_upload_nodes(h, n_x, x2, p2))) return rc;
    if ((rc = ensure_panels(h, np))) return rc;
    if ((rc = th_ensure_eig(h, q))) return rc;   // the spectral path where th_chi0 takes it
    th_green_args a{};
    a.mode = mode; a.p0 = p0; a.np = np; a.q[0] = q[0]; a.q[1] = q[1]; a.q[2] = q[2];
    a.omega = omega; a.Ef = Ef; a.eta = eta; a.etap = etap; a.prec = prec; a.panel = h->d_pA;
    if (mode >= TH_G_REAL_A) {
        a.kmajor_re = h->pair_order;
        a.nx = n_x;
    }
    if ((rc = th_launch_green(h, a))) return rc;
    const size_t n = (size_t)np * h->nA2 * 2 * h->Mpad;
    if (prec == TH_PREC_FP64) {
        TH_HIP(hipMemcpyAsync(out, h->d_pA, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    } else {
        std::vector<float> tmp(n);
        TH_HIP(hipMemcpyAsync(tmp.data(), h->d_pA, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
        TH_HIP(hipStreamSynchronize(h->stream));
        for (size_t e = 0; e < n; ++e) out[e] = tmp[e];
    }
    return th_check_status(h);
}

int thx_last_split(th_handle *h) { return h ? h->last_presplit : 0; }
int thx_last_refined(th_handle *h) { return h ? h->last_refined : 0; }
int thx_last_spectral(th_handle *h) { return h ? h->last_spectral : 0; }

int thx_eigenpairs(th_handle *h, const double *q, double *L_out, double *V_out)
{
    if (!h || bad(q) || bad(L_out) || bad(V_out)) { th_set_error("thx_eigenpairs: bad arguments"); return TH_E_ARG; }
    int rc = th_ensure_eig(h, q);
    if (rc) return rc;
    const bool qz = q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0;
    const int which = qz ? 0 : 1;
    if (!h->eig_valid[which]) { th_set_error("thx_eigenpairs: no eigenpairs (spectral path off)"); return TH_E_STATE; }
    const size_t nk = (size_t)h->n_kuniq, d = (size_t)h->d;
    TH_HIP(hipMemcpyAsync(L_out, h->eig[which].L, sizeof(double) * nk * d, hipMemcpyDeviceToHost, h->stream));
    TH_HIP(hipMemcpyAsync(V_out, h->eig[which].V, sizeof(double2) * nk * d * d, hipMemcpyDeviceToHost, h->stream));
    TH_HIP(hipStreamSynchronize(h->stream));
    return TH_OK;
}

int thx_contract(th_handle *h, const double *A, const double *B, long np, int prec, double *acc_out)
{
    if (!h || np < 1 || bad(A) || bad(B) || bad(acc_out)) { th_set_error("thx_contract: bad arguments"); return TH_E_ARG; }
    int rc = ensure_panels(h, np);
    if (rc) return rc;
    if ((rc = th_ensure_part(h, th_contract_splits(h, np, prec)))) return rc;
    const size_t n = (size_t)np * h->nA2 * 2 * h->Mpad;
    if (prec == TH_PREC_FP64) {
        TH_HIP(hipMemcpy(h->d_pA, A, sizeof(double) * n, hipMemcpyHostToDevice));
        TH_HIP(hipMemcpy(h->d_pB, B, sizeof(double) * n, hipMemcpyHostToDevice));
    } else {
        std::vector<float> a(n), b(n);
        for (size_t e = 0; e < n; ++e) { a[e] = (float)A[e]; b[e] = (float)B[e]; }
        TH_HIP(hipMemcpy(h->d_pA, a.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        TH_HIP(hipMemcpy(h->d_pB, b.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    const size_t nacc = (size_t)h->nA2 * 2 * h->Mpad * h->Mpad;
    if ((rc = th_launch_zero(h, h->d_acc, nacc))) return rc;
    // scales of the half-precision contraction from the panels' largest entries
    th_contract_scale sc;
    if (prec != TH_PREC_FP64) {
        float ma = 0.0f, mb = 0.0f;
        for (size_t e = 0; e < n; ++e) {
            ma = std::max(ma, std::fabs((float)A[e]));
            mb = std::max(mb, std::fabs((float)B[e]));
        }
        sc.eA = th_pow2_scale(ma);
        sc.eB = th_pow2_scale(mb);
    }
    if ((rc = th_launch_contract(h, h->d_pA, h->d_pB, np, prec, 0, &sc))) return rc;
    TH_HIP(hipMemcpyAsync(acc_out, h->d_acc, sizeof(double) * nacc, hipMemcpyDeviceToHost, h->stream));
    TH_HIP(hipStreamSynchronize(h->stream));
    return TH_OK;
}

}  // extern "C"
