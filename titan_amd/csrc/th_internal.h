// th_internal.h -- shared definitions of the gfx950 kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/titan_hip.h"

// Panel geometry: the G*G contraction runs one complex GEMM per atom pair
// (i,j) whose rows and columns are the flattened (2 nOrb) x (2 nOrb) atom
// blocks, M = (2 nOrb)^2 (324 for s,p,d), padded to MPAD, a multiple of the
// 112-wide output tile.
constexpr int TH_TILE = 112;        // output tile edge (7 MFMA blocks of 16)
constexpr int TH_KC = 8;            // pairs per LDS stage of the GEMM
constexpr int TH_RUN = 16;          // pairs of one k-point per workgroup of the run-per-k spectral kernel
constexpr int TH_GEMM_WAVES = 7;    // one wave per 16-row block of the tile
constexpr int TH_MAX_NORB = 9;
constexpr int TH_MAX_D_LDS = 64;    // largest d of the generic LDS-resident inversion
constexpr int TH_MAX_D = 288;       // largest d of any inversion kernel (288: block elimination)
constexpr int TH_MAX_NA = 32;

inline int th_mpad(int M) { return ((M + TH_TILE - 1) / TH_TILE) * TH_TILE; }

// Node ie and k-point ik of real-axis pair p in the grouped k-major order
// (the run-per-k Green kernels need TH_RUN consecutive pairs of one k-point):
// the nx nodes split into blocks of TH_RUN, the last one shorter; block b
// holds its nodes for every k-point, k-point by k-point, so
// p = b TH_RUN nk + ik len(b) + (ie - b TH_RUN).  Pair counts stay below 2^31.
__host__ __device__ inline void th_real_grouped(unsigned p, unsigned nk, unsigned nx, unsigned &ie, unsigned &ik)
{
    const unsigned full = (unsigned)TH_RUN * nk, nb = (nx + TH_RUN - 1) / TH_RUN;
    const unsigned bq = p / full, b = bq < nb ? bq : nb - 1;
    const unsigned off = p - b * full, rest = nx - (unsigned)TH_RUN * b;
    const unsigned bl = rest < (unsigned)TH_RUN ? rest : (unsigned)TH_RUN;
    ik = off / bl;
    ie = (unsigned)TH_RUN * b + (off - ik * bl);
}

// eigenpairs of H(k + q) for the unique k-points of the work lists
struct th_eig_set {
    double2 *V = nullptr;        // [n_kuniq][d][d] row-major, V[r][m] = component r of eigenvector m
    double *L = nullptr;         // [n_kuniq][d]
    // V as the fp16 matrix-core kernels read it (d = 72 and 144, else
    // nullptr): per k-point four planes (Re hi, Re lo, Im hi, Im lo) of the
    // two-term split of 2^14 V, [4][RP][KP] halves with the rows padded to RP =
    // 16 ceil(d/16) and the eigen-indices to KP = 32 ceil(d/32), zero there
    void *S = nullptr;
    long cap = 0;                // k-points the buffers hold
    double q[3] = {0, 0, 0};
};

struct th_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of the chi0 pipeline: the inversions of batch b + 1 run
    // there while the MFMA contraction of batch b runs on 'stream'
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_green[2] = {nullptr, nullptr}, ev_contract[2] = {nullptr, nullptr};
    int overlap = 1;             // pipeline the inversions and the contraction on two streams
    void *d_pA2 = nullptr, *d_pB2 = nullptr;   // second panel slot of the pipeline
    int nA = 0, nOrb = 0, bs = 0, d = 0, D = 0, nhop = 0, M = 0, Mpad = 0, nA2 = 0;
    // model
    double *d_R = nullptr;       // 3*nhop
    int *d_jb = nullptr, *d_ib = nullptr;
    double2 *d_t = nullptr;      // nhop * nOrb*nOrb (column-major blocks), sorted by (jb, ib)
    int *d_hoff = nullptr;       // nA*nA+1 offsets: hops of atom pair (jb, ib) are hoff[jb*nA+ib] ..
    double2 *d_h0 = nullptr;     // d*d column-major
    // work lists
    long n_im = 0, nk_re = 0;
    double *d_kim = nullptr, *d_yim = nullptr, *d_wim = nullptr;
    double *d_kre = nullptr, *d_wre = nullptr;
    double *d_x2 = nullptr, *d_p2 = nullptr;  // real-axis nodes of the current omega
    int nx_cap = 0;
    // work space
    long batch = 0;              // pairs per batch
    void *d_pA = nullptr, *d_pB = nullptr;    // panels [batch][nA2][2][Mpad]
    size_t panel_bytes = 0;
    int splits = 0;
    double *d_part = nullptr;    // [part_cap][nA2][2][Mpad][Mpad]
    int part_cap = 0;            // splits the partial buffer holds
    double *d_acc = nullptr;     // [2 axes][nA2][2][Mpad][Mpad]
    double2 *d_chi0 = nullptr;   // D*D
    double2 *d_chi = nullptr;    // D*D
    // omega-independent imaginary-axis B panels (G(k, E_F + i(y+etap)))
    void *d_bcache = nullptr;
    size_t bcache_bytes = 0;     // size of the d_bcache allocation
    int bcache_prec = -1;
    double bcache_key[2] = {0, 0};   // (Ef, etap)
    long bcache_bp = 0;              // stored in summation order per batch of this many pairs (0: pair order)
    int bcache_presplit = 0;         // rows hold fp16 splits (th_green_args::presplit)
    int bcache_pexp = 0;             // their exponent
    // Dyson work space
    int *d_sup = nullptr;        // support of U
    double2 *d_U = nullptr, *d_W = nullptr, *d_V = nullptr, *d_Z = nullptr, *d_X = nullptr;
    int sup_cap = 0;
    // timing
    hipEvent_t ev[8] = {};
    double t_ms[5] = {0, 0, 0, 0, 0};
    long n_launch[4] = {0, 0, 0, 0};
    int kernel_variant = 0;      // 0 default kernels, 1..3 alternates (tuning / comparison)
    int use_graph = 0;           // th_chi_sweep replays one captured hipGraph per omega
    int contract_3m = 0;         // contraction with three real products per complex one
    // spectral Green functions (Hermitian H, d = 18, 36, 54)
    int hermitian = 0;           // H(k) is Hermitian for every k (checked at th_create)
    int spectral = 1;            // use them when possible
    int spectral_chunk = 1;      // spectral G for runs of pairs per k-point (V loaded once per run)
    int contract_f16 = 1;        // mixed: contraction on fp16 MFMA with two-term splits of the FP32 panels
    void *rb = nullptr;          // rocblas handle of the dense eigen-decompositions (d > 64), made on first use
    int current_general = 0;     // th_current_response: three-stage form for every d (1) or only where LDS cannot hold a pair (0)
    int contract_tiles = 1;      // split-panel contraction: 8 waves with 2-D register tiles (1) or 7 strips (0)
    int presplit = 1;            // mixed spectral: Green function kernels write the fp16 splits themselves
    int last_presplit = 0;       // the last chi0 did
    long wave_pairs = -1;        // pairs of one full wave of fp16 spectral workgroups (0: none; -1: not yet known)
    int green_f16 = 1;           // mixed: spectral products on fp16 MFMA with two-term splits
    int pair_order = 1;          // imaginary list sorted by k-point, real axis k-major (runs per k-point)
    long n_kuniq = 0;            // unique k-points of the imaginary list and the real mesh
    double *d_kuniq = nullptr;   // [n_kuniq][3]
    int *d_kid_im = nullptr, *d_kid_re = nullptr;   // pair / mesh point -> unique k-point
    th_eig_set eig[2];           // [0]: q = 0, [1]: the last q != 0
    int eig_valid[2] = {0, 0};
    std::vector<int> kid_im_local;        // imaginary point -> index in kim_uniq
    long list_version = 0;                // bumped whenever the imaginary list or the real mesh is set
    std::vector<int> kid_im_host, kid_re_host;   // d_kid_im, d_kid_re on the host
    double2 *d_wk = nullptr;              // adjoint: per unique k-point and side, the eigenbasis Hhat sums
    size_t wk_bytes = 0;
    long last_blkmax = 0;                 // workgroup maxima the last th_launch_green wrote (0: none)
    int adj_bcache = 0;                   // 1: the adjoint reads the imaginary-axis B panels from the forward's cache
    // the adjoint's segments (runs of pairs on one k-point within a batch) per
    // axis, for the key {list_version, pairs per batch, real-axis nodes}
    std::vector<int> adj_seg[2];
    // per batch its segments (index in the batch) in rounds of distinct
    // k-points, and the rounds' offsets into adj_sord
    std::vector<int> adj_sord[2];
    std::vector<std::vector<long>> adj_roff[2];
    long adj_seg_key[3] = {-1, -1, -1};
    std::vector<double> kim_uniq, kre_host;
    std::vector<double> yim_host, wim_host;   // imaginary list in stored order
    double wre_absmax = 0.0;     // max |w| of the real k-mesh
    double im_bound_key[2] = {-1.0, -1.0};    // (eta, etap) of the cached bounds below
    double im_bound[2] = {0.0, 0.0};          // max |w|/(y + eta), max 1/(y + etap) over the imaginary list
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // [omega == 0, omega != 0]
    int graph_prec[2] = {-1, -1};
    double *d_omega = nullptr;   // omega of the current replay
    double *d_norms = nullptr;   // Dyson norms of the current replay
    int batch_mb = 512;          // panel budget per batch (MB)
    long cache_mb = 0;           // cap of the omega-independent caches (B panels, eigenpairs), MB; 0: none
    // adjoint work space and the caller's hop order (hops are stored sorted)
    std::vector<int> hop_order;  // stored hop n came from caller's entry hop_order[n]
    void *d_adj = nullptr;
    void *d_pamax = nullptr;     // two words: the largest |X| and |That| of the fp16 panel adjoint
    void *d_pays = nullptr;      // That of the fp16 panel adjoint split into fp16 planes (both transpositions)
    const double *pa_that = nullptr;   // the That d_pays holds
    size_t adj_bytes = 0;
    void *d_blk = nullptr;       // work space of the block elimination (d = 288)
    size_t blk_bytes = 0;
    // imaginary pairs and real-mesh points of each unique k-point (CSR, for
    // the per-k-point FP64 block of chi0 on the support of U)
    int *d_kim_ptr = nullptr, *d_kim_idx = nullptr, *d_kre_ptr = nullptr, *d_kre_idx = nullptr;
    // FP64 chi0[S,S] of the mixed path (th_refine.hip)
    int refine = 1;              // recompute chi0 on the support of U in FP64 (mixed precision)
    int last_refined = 0;        // the last chi0 did
    int last_spectral = 0;       // the last chi0 took its Green functions from eigenpairs
    std::vector<int> ref_S;      // support of the declared U (chi indices, ascending)
    std::vector<int> ref_built_S;   // support the device tables below were built for
    long ref_built_nk = -1;
    int ref_ready = 0, ref_nR = 0, ref_ntile = 0, ref_nxc = 1, ref_kpc = 1, ref_nchunk = 0;
    int ref_mfma = 0;            // the support fits the matrix-core form (at most 32 indices)
    size_t ref_smem = 0, ref_part_bytes = 0;
    int *d_ref_int = nullptr;    // R, sx, sxp, sy, syp, tiles, S sorted
    void *d_ref_part = nullptr;  // per-chunk partial blocks
};

// error plumbing
void th_set_error(const std::string &msg);
#define TH_HIP(call)                                                               \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            th_set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
            return TH_E_HIP;                                                       \
        }                                                                          \
    } while (0)

// Stage timing with HIP events on the handle's stream
struct stage_timer {
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
    std::vector<hipStream_t> st;
    th_handle *h;
    explicit stage_timer(th_handle *hh) : h(hh) {}
    int begin(int stage, hipStream_t s = nullptr)
    {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
        if (!s) s = h->stream;
        hipEventRecord(a, s);
        ev.push_back({stage, {a, b}});
        st.push_back(s);
        return (int)ev.size() - 1;
    }
    void end(int id)
    {
        if (id >= 0) hipEventRecord(ev[id].second.second, st[id]);
    }
    void collect()
    {
        for (int s = 0; s < 5; ++s) h->t_ms[s] = 0.0;
        for (auto &e : ev) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess) h->t_ms[e.first] += ms;
        }
    }
    ~stage_timer()
    {
        for (auto &e : ev) {
            hipEventDestroy(e.second.first);
            hipEventDestroy(e.second.second);
        }
    }
};

// shared helpers of the C ABI (th_abi.hip)
int *th_status_word(th_handle *h);
int th_reset_status(th_handle *h);
int th_check_status(th_handle *h);
int th_ensure_panels(th_handle *h, long pairs);
int th_upload_nodes(th_handle *h, int n_x, const double *x2, const double *p2);
int th_ensure_eig(th_handle *h, const double *q);

// ---- kernel launchers (th_green.hip) ----
// Modes of the green-function kernel: which matrix, and how it is written
// into the GEMM panel.
enum th_green_mode {
    TH_G_IMAG_A = 0,   // A = w * Gq,                 Gq = G(k+q, Ef+omega + i(y+eta))
    TH_G_IMAG_B = 1,   // B = G(j,i) blocks,          G  = G(k, Ef + i(y+etap))
    TH_G_REAL_A = 2,   // A = -i w (Gq - Gq^dagger),  Gq = G(k+q, x+omega + i eta)
    TH_G_REAL_B = 3,   // B = G^dagger(j,i) blocks,   G  = G(k, x + i etap)
};

struct th_green_args {
    int mode;
    long p0, np;          // pair range of this batch (global pair index p0 .. p0+np-1)
    double q[3];
    double omega, Ef, eta, etap;
    int prec;             // TH_PREC_*
    void *panel;          // output panel, [np][nA2][2][Mpad] of double or float
    double2 *raw = nullptr;   // optional: G itself, [np][d][d] row-major FP64
    const double *omega_p = nullptr;   // optional: omega in device memory (graph replays)
    hipStream_t stream = nullptr;      // optional: launch stream (default: the handle's)
    int kmajor_re = 0, nx = 0;         // real axis: pairs grouped k-major (th_real_grouped) instead of node-major
    // panel rows in the contraction's summation order (with pair_order: the
    // runs of TH_RUN pairs per k-point stored transposed, see panel_row)
    int sum_order = 0;
    // half-precision panels: rows of [nA2][re|im][Mpad] words, each the fp16
    // pair (hi, lo) of the two-term split of x 2^pexp, hi in the low half (the
    // half-precision spectral kernel only); on
    // the real axis with pareal > 0 the exponent is taken on the device from
    // pareal max |p2[0 .. pnx)| instead
    int presplit = 0, pexp = 0, pnx = 0;
    double pareal = 0.0;
    // optional, FP32 panels: each workgroup's largest |entry| of what it wrote,
    // [np / TH_RUN rounded up]; kernels that support it set th_handle::last_blkmax
    float *blkmax = nullptr;
};
int th_launch_green(th_handle *h, const th_green_args &a);
// The Green function kernels of both sides of chi0(q) can write split
// half-precision panels (th_green_args::presplit)
bool th_green_presplit_ok(const th_handle *h, const double *q);
// the eigenpairs of one side's k-points (qside: k + q), when ready for q
const th_eig_set *th_side_eig(const th_handle *h, bool qside, const double *q);
// Pairs that fill every resident workgroup of the fp16 spectral Green kernel
// once (TH_RUN pairs per workgroup x compute units x workgroups per unit), 0
// when the handle's d has no such kernel; batches of whole multiples leave no
// part-empty last wave
long th_green_wave_pairs(th_handle *h);
int th_launch_eig(th_handle *h, int which, const double *q);
// d with spectral Green function kernels (and an eigen-decomposition for them)
bool th_spectral_d(int d);
// halves per k-point of th_eig_set::S for this d (0: not kept)
size_t th_eig_split_halves(int d);
// releases the library handle of the dense eigen-decompositions
void th_eig_release(th_handle *h);

// Row of global pair b in panels stored per batch of bpc pairs with each
// batch's rows in the contraction's summation order (the runs of TH_RUN pairs
// per k-point transposed, as the Green function kernels write them with
// sum_order); bpc = 0: row b.  pg0 is added to batch-local pair indices.
struct th_rowmap {
    long pg0 = 0, bpc = 0, ntot = 0;
};
__host__ __device__ inline long th_map_row(long b, const th_rowmap &m)
{
    if (m.bpc <= 0) return b;
    const long cb = b / m.bpc, l = b - cb * m.bpc;
    const long nb = (m.ntot - cb * m.bpc) < m.bpc ? (m.ntot - cb * m.bpc) : m.bpc;
    const long n = nb / TH_RUN;
    return cb * m.bpc + (l < TH_RUN * n ? (l % TH_RUN) * n + l / TH_RUN : l);
}

// ---- contraction (th_gemm.hip) ----
// Power-of-two scales of the two FP32 panels for the half-precision
// contraction: every |entry| 2^e of a panel stays below 2^14.  On the real axis
// the A exponent is taken on the device from areal max_x |p2[x]| (the node
// weights of the current frequency, so that graph replays see new ones).
struct th_contract_scale {
    int eA = 0, eB = 0;
    const double *p2 = nullptr;
    int n_x = 0;
    double areal = 0.0;
    bool presplit = false;   // panels already split into fp16 hi + lo (th_green_args::presplit)
};
// Exponent of the real-axis A panel: bound = areal max |p2[x]|, e = 14 - E
// with bound = m 2^E, m in [0.5, 1) (the producer and the contraction use
// this one function so that both see the same e)
__device__ inline int th_real_exp(const double *p2, int n_x, double areal)
{
    double m = 0.0;
    for (int x = 0; x < n_x; ++x) m = fmax(m, fabs(p2[x]));
    const double bound = areal * m;
    const int ex = (int)((__double_as_longlong(bound) >> 52) & 0x7ff) - 1022;
    return bound > 0.0 ? min(100, max(-100, 14 - ex)) : 0;
}
// Exponent e with bound 2^e < 2^14 (0 for a bound that is not positive and finite)
int th_pow2_scale(double bound);
// acc[axis] += sum over the batch of A (x) B, one complex GEMM per atom pair.
// sc: the panels' scales (FP32 only); without them the FP32 kernel runs.
int th_launch_contract(th_handle *h, const void *pA, const void *pB, long np, int prec, int axis,
                       const th_contract_scale *sc = nullptr);
// K-splits the contraction uses for a batch of np pairs: the tuned value,
// raised in FP32 so that no partial sum runs over more than 8192 pairs
int th_contract_splits(const th_handle *h, long np, int prec);
int th_ensure_part(th_handle *h, int splits);
int th_launch_zero(th_handle *h, double *p, size_t n);
// chi0 = (T_imag + T_imag^dagger + T_real) / 2pi in the chi basis.
int th_launch_finalize(th_handle *h, double2 *chi0);
// Adjoint of the contraction for one panel: out[p] = conj(panel[p]) . That^T
// (trans = 1) or conj(panel[p]) . That (trans = 0) per atom pair, FP64 out
// [np][nA2][2][Mpad]; That is [nA2][2][Mpad][Mpad] FP64.
// split_e != INT_MIN: the panel holds split (hi, lo) words of x 2^split_e (FP32 only)
// blkmax / nblkmax: the panel's largest |entry| as per-workgroup maxima from
// the Green kernel that wrote it (instead of a pass over the panel)
int th_launch_panel_adjoint(th_handle *h, const void *panel, int prec, const double *That, long np, int trans,
                            double *out, const th_rowmap &rows = th_rowmap(), int split_e = INT_MIN,
                            const float *blkmax = nullptr, long nblkmax = 0);
// For the fp16 panel adjoint (mixed precision): That of one axis split once
// into the fp16 layouts of both transpositions before the axis's batches
// (its largest |That| kept on the device); later launches with this That use
// them.  FP64, or contract_f16 off: nothing to do.
int th_panel_adjoint_prepare(th_handle *h, int prec, const double *That);

// ---- FP64 chi0 on the support of U (th_refine.hip) ----
constexpr int TH_REFINE_MAX_S = 1024;   // larger supports are not refined
// Support of U (D x D host, column-major; NULL clears it)
int th_refine_set_support(th_handle *h, const double2 *U);
// Host preparation (tables, partial buffer); *on = the block will be refined
int th_refine_prepare(th_handle *h, const double *q, int prec, bool *on);
// Launches only (capturable): chi0[S,S] overwritten by the FP64 block
int th_refine_launch(th_handle *h, const double *q, double omega, const double *omega_p, double Ef, double eta,
                     double etap, int n_x, bool real_axis, double2 *chi0);

// ---- Dyson (th_dyson.hip) ----
// device scratch of the allocation-free Dyson entry point, per device
constexpr size_t TH_NOALLOC_SCRATCH = (size_t)64 << 20;
int th_noalloc_init(int device);   // once per device, that device current
int th_dyson_noalloc_impl(th_handle *h, const double2 *chi0_dev, const double2 *U_host, double2 *chi_dev,
                          double *cond);
int th_dyson_device(th_handle *h, const double2 *chi0, const double2 *U_host_colmajor,
                    double2 *chi, double *cond);
// Graph-friendly split of the Dyson step: set-up from U (host work, uploads,
// allocations; returns the support size r, 0 if U = 0), then kernels only,
// norms ||W||_1, ||W^-1||_1 left in norms_dev[0..1].
int th_dyson_prepare(th_handle *h, const double2 *U_host_colmajor, int *r_out);
int th_dyson_launch(th_handle *h, int r, const double2 *chi0, double2 *chi, double *norms_dev);
int th_launch_trace(th_handle *h, const double2 *chi, double2 *schi_dev);
