// th_gemm.hip -- the chi0 contraction on the MFMA units of gfx950.
//
// eintshechi's six-deep loop (eintshechi.F90:83-105) accumulates, for every
// atom pair (i,j), sum_p w_p Gq_p(i,j)[ab] G_p(j,i)[cd] over the (k,E)
// pairs p.  With the atom blocks flattened to vectors of length
// M = (2 nOrb)^2 this is one complex GEMM per atom pair,
//     T(i,j) = sum_p A_p B_p^T,   A_p = vec(w_p Gq_p(i,j)), B_p = vec(G_p(j,i)),
// with K = number of pairs.  The second term of the reference,
// conj(G(i,j) Gq(j,i)), is T^dagger in the chi index space (applied in the
// finalize kernel), so it costs no extra MFMA work.  The real-axis integral
// (eintshechi.F90:143-165) is the same GEMM with A = -i w (Gq - Gq^dagger),
// B = G^dagger, accumulated separately because it is not symmetrised.
//
// Kernel layout: a workgroup owns one 112 x 112 output tile of one atom pair
// for one K-split.  Seven waves each hold a 16 x 112 strip as 7 MFMA blocks
// (re and im accumulators, 4 real MFMAs per complex block step).  Panels
// stream through LDS in stages of TH_KC pairs, double buffered with a
// register prefetch.  Partial tiles go to a split buffer and are summed in
// a fixed order (bitwise reproducible).
#include <algorithm>
#include <climits>
#include <cmath>

#include "th_internal.h"

namespace {

typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));

template <typename P> struct mfma_t;
// Output layout of the 16x16 block: lane l, register q holds row
// row(l >> 4, q) and column l & 15 (measured on gfx950 with thx_contract).
template <> struct mfma_t<double> {
    typedef dv4 acc;
    static __device__ inline acc mac(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int g, int q) { return g + 4 * q; }
};
template <> struct mfma_t<float> {
    typedef fv4 acc;
    static __device__ inline acc mac(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int g, int q) { return 4 * g + q; }
};

constexpr int NTH = TH_GEMM_WAVES * 64;   // 448 threads

// 16-byte vector of panel elements
template <typename P> struct vec16;
template <> struct vec16<double> { typedef double2 type; static constexpr int n = 2; };
template <> struct vec16<float> { typedef float4 type; static constexpr int n = 4; };

// K3M = true: three real products per complex block step (Karatsuba / 3M:
// P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); re = P1 - P2,
// im = P3 - P1 - P2) instead of four, 25 % less MFMA work.
template <typename P, bool K3M>
__global__ __launch_bounds__(NTH) void contract_kernel(const P *__restrict__ A, const P *__restrict__ B,
                                                       long np, int Mpad, int nA2, int ntile, int splits,
                                                       long kper, double *__restrict__ part)
{
    typedef typename vec16<P>::type V;
    constexpr int VN = vec16<P>::n;
    constexpr int ROWV = TH_TILE / VN;                 // vectors per 112-wide row
    constexpr int STAGE_V = TH_KC * 2 * ROWV;          // vectors per matrix per stage
    constexpr int PER_T = (STAGE_V + NTH - 1) / NTH;   // vectors per thread per matrix
    __shared__ __align__(16) P As[2][TH_KC][2][TH_TILE];
    __shared__ __align__(16) P Bs[2][TH_KC][2][TH_TILE];

    // XCD-aware order: consecutive logical ids (the tiles of one split) land
    // on the same XCD, whose L2 then serves their shared panel rows.
    // (workgroups are dispatched round-robin over the 8 XCDs: b % 8 is the XCD;
    // the tail beyond a multiple of 8 keeps its own index)
    const int nwg = gridDim.x;
    const int b = blockIdx.x;
    const int n8 = nwg & ~7;
    const int lid = (b < n8) ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
    const int tile = lid % (ntile * ntile);
    const int rest = lid / (ntile * ntile);
    const int split = rest % splits;
    const int ij = rest / splits;
    const int tm = tile / ntile, tn = tile - tm * ntile;
    const int r0 = tm * TH_TILE, c0 = tn * TH_TILE;
    const long kb = (long)split * kper;
    const long ke = min(np, kb + kper);

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const size_t pstride = (size_t)nA2 * 2 * Mpad;       // elements per pair in a panel
    const P *Ab = A + (size_t)ij * 2 * Mpad + r0;
    const P *Bb = B + (size_t)ij * 2 * Mpad + c0;

    typename mfma_t<P>::acc accr[7], acci[7], acc3[K3M ? 7 : 1];
    for (int j = 0; j < 7; ++j)
        for (int q = 0; q < 4; ++q) { accr[j][q] = 0; acci[j][q] = 0; }
    for (int j = 0; j < (K3M ? 7 : 1); ++j)
        for (int q = 0; q < 4; ++q) acc3[j][q] = 0;

    V ra[PER_T], rb[PER_T];
    auto gload = [&](long k0) {
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            int v = tid + u * NTH;
            V za, zb;
            for (int q = 0; q < VN; ++q) { reinterpret_cast<P *>(&za)[q] = 0; reinterpret_cast<P *>(&zb)[q] = 0; }
            if (v < STAGE_V) {
                int kk = v / (2 * ROWV), rem = v - kk * 2 * ROWV, comp = rem / ROWV, col = rem - comp * ROWV;
                long kp = k0 + kk;
                if (kp < ke) {
                    const size_t row = (size_t)kp;
                    za = *reinterpret_cast<const V *>(Ab + row * pstride + (size_t)comp * Mpad + col * VN);
                    zb = *reinterpret_cast<const V *>(Bb + row * pstride + (size_t)comp * Mpad + col * VN);
                }
            }
            ra[u] = za;
            rb[u] = zb;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            int v = tid + u * NTH;
            if (v < STAGE_V) {
                int kk = v / (2 * ROWV), rem = v - kk * 2 * ROWV, comp = rem / ROWV, col = rem - comp * ROWV;
                *reinterpret_cast<V *>(&As[buf][kk][comp][col * VN]) = ra[u];
                *reinterpret_cast<V *>(&Bs[buf][kk][comp][col * VN]) = rb[u];
            }
        }
    };

    const long nstage = (ke > kb) ? (ke - kb + TH_KC - 1) / TH_KC : 0;
    if (nstage > 0) gload(kb);
    for (long s = 0; s < nstage; ++s) {
        const int buf = (int)(s & 1);
        sstore(buf);
        __syncthreads();
        if (s + 1 < nstage) gload(kb + (s + 1) * TH_KC);
#pragma unroll
        for (int kk = 0; kk < TH_KC; kk += 4) {
            const int pr = kk + (lane >> 4);
            const int m = wave * 16 + (lane & 15);
            const P ar = As[buf][pr][0][m], ai = As[buf][pr][1][m];
            if constexpr (K3M) {
                const P as = ar + ai;
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const int n = j * 16 + (lane & 15);
                    const P br = Bs[buf][pr][0][n], bi = Bs[buf][pr][1][n];
                    accr[j] = mfma_t<P>::mac(ar, br, accr[j]);        // P1
                    acci[j] = mfma_t<P>::mac(ai, bi, acci[j]);        // P2
                    acc3[j] = mfma_t<P>::mac(as, br + bi, acc3[j]);   // P3
                }
            } else {
                const P nai = -ai;
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const int n = j * 16 + (lane & 15);
                    const P br = Bs[buf][pr][0][n], bi = Bs[buf][pr][1][n];
                    accr[j] = mfma_t<P>::mac(ar, br, accr[j]);
                    accr[j] = mfma_t<P>::mac(nai, bi, accr[j]);
                    acci[j] = mfma_t<P>::mac(ar, bi, acci[j]);
                    acci[j] = mfma_t<P>::mac(ai, br, acci[j]);
                }
            }
        }
    }
    // partial tile -> part[split][ij][comp][Mpad][Mpad]
    double *pr_ = part + (((size_t)split * nA2 + ij) * 2 + 0) * Mpad * Mpad;
    double *pi_ = pr_ + (size_t)Mpad * Mpad;
#pragma unroll
    for (int j = 0; j < 7; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + wave * 16 + mfma_t<P>::row(lane >> 4, q);
            const int col = c0 + j * 16 + (lane & 15);
            double re, im;
            if constexpr (K3M) {
                const double p1 = (double)accr[j][q], p2 = (double)acci[j][q], p3 = (double)acc3[j][q];
                re = p1 - p2;
                im = p3 - p1 - p2;
            } else {
                re = (double)accr[j][q];
                im = (double)acci[j][q];
            }
            pr_[(size_t)row * Mpad + col] = re;
            pi_[(size_t)row * Mpad + col] = im;
        }
}

// The contraction for FP32 panels on the half-precision matrix cores
// (v_mfma_f32_16x16x32_f16: 32 pairs per instruction at 16x the FP32 MFMA
// rate).  Each FP32 value x is split as x 2^e = hi + lo with hi = fp16(x 2^e)
// and lo = fp16(x 2^e - hi) (22 significant bits together), and each real
// product is hi.hi + hi.lo + lo.hi accumulated in FP32 (the dropped lo.lo is
// 2^-22 of the product): three half-precision MFMAs instead of eight FP32 ones
// per 32 pairs.  e is one power of two per operand and launch, from the
// resolvent bound |G| <= 1/Im z of a Hermitian H (imaginary axis: |w G_q| <=
// |w|/(y + eta), |G| <= 1/(y + eta'); real axis: |w (G_q - G_q^H)| <=
// 2|w|/eta, |G^H| <= 1/eta'), so that every scaled value stays below 2^14
// (fp16 holds 65504); a value far below the bound keeps an absolute error
// below 2^-39 of it (half the fp16 subnormal spacing in lo), so the bound may
// be loose by orders of magnitude.  The partial sums leave scaled back by
// 2^-(eA+eB) in FP64.  H must be Hermitian (else the FP32 kernel runs).
//
// LDS holds two stages (double buffer) of both operands as [operand][re|im]
// [hi|lo][112 rows][32 pairs + 8] fp16 (rows padded to 80 B so that the
// 16-byte fragment reads of 8 lanes fall on distinct banks): 143,360 B, one
// workgroup per CU.  Stage s + 1 is split into the free buffer in the same
// basic block as the MFMAs of stage s (so that the scheduler interleaves the
// conversion with the matrix work), the loads run two stages ahead, one
// barrier per stage.  Lane l of wave w reads the 8 pairs 8 (l >> 4) .. +7 of
// row 16 w + (l & 15) (A) or of column row 16 j + (l & 15) (B); results land
// in rows 4 (l >> 4) + q.
constexpr int HK = 32;          // pairs per stage (K of one MFMA)
constexpr int HST = HK + 8;     // LDS row stride in halves
typedef _Float16 hv8 __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(NTH) void contract_h2_kernel(const float *__restrict__ A, const float *__restrict__ B,
                                                          long np, int Mpad, int nA2, int ntile, int splits,
                                                          long kper, int eA, int eB,
                                                          const double *__restrict__ p2, int n_x, double areal,
                                                          double *__restrict__ part)
{
    // [buffer][operand][re|im][hi|lo][row * HST + pair]
    __shared__ __align__(16) _Float16 Ls[2][2][2][2][TH_TILE * HST];
    const int nwg = gridDim.x;
    const int b = blockIdx.x;
    const int n8 = nwg & ~7;
    const int lid = (b < n8) ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
    const int tile = lid % (ntile * ntile);
    const int rest = lid / (ntile * ntile);
    const int split = rest % splits;
    const int ij = rest / splits;
    const int tm = tile / ntile, tn = tile - tm * ntile;
    const int r0 = tm * TH_TILE, c0 = tn * TH_TILE;
    // pair indices of a batch and byte offsets within its panels fit 32 bits
    // (checked on the host)
    const unsigned kb = (unsigned)(split * kper);
    const unsigned ke = (unsigned)min(np, (long)kb + kper);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // raw buffer loads over the split's rows: an offset at or past row ke
    // reads 0 without touching memory (no select waits on loaded values).
    // The B resource starts (c0 - r0) floats off B so that one offset
    // serves both operands; the imaginary halves sit Mpad floats on (scalar
    // offset).  Panel bytes + 96 rows < 2^31 (checked on the host).
    const unsigned ps4 = (unsigned)nA2 * 2u * (unsigned)Mpad * 4u;
    const int nrec = (int)(ke * ps4);
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc((void *)A, (short)0, nrec, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(reinterpret_cast<const char *>(B) + (long)(c0 - r0) * 4), (short)0, nrec, 0x00020000);
    // real axis: |A| <= 2 max|p2| max|w_k| / eta = areal max|p2| (omega-dependent, from device memory)
    if (p2) eA = th_real_exp(p2, n_x, areal);
    const float scA = ldexpf(1.0f, eA), scB = ldexpf(1.0f, eB);

    // staging: thread tid holds column e = tid % 112 of the pairs 8 g .. 8 g + 7
    // (g = tid / 112) of a stage, both components, both operands
    const int se = tid % TH_TILE, sg = tid / TH_TILE;
    const unsigned base = ((unsigned)ij * 2u * (unsigned)Mpad + (unsigned)(r0 + se)) * 4u;
    const int imoff = Mpad * 4;
    float va[2][8], vb[2][8];
    auto gload = [&](unsigned k0) {   // issue the loads of a stage
        const unsigned v0 = base + (k0 + 8u * (unsigned)sg) * ps4;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int vo = (int)(v0 + (unsigned)i * ps4);
            va[0][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rA, vo, 0, 0));
            va[1][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rA, vo, imoff, 0));
            vb[0][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rB, vo, 0, 0));
            vb[1][i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rB, vo, imoff, 0));
        }
    };
    // x 2^e = hi + lo: hi = fp16(x 2^e), lo = fp16(x 2^e - hi) with one fused
    // multiply-add on the fp32 value and the fp16 hi
    auto split_store = [&](int buf) {
#pragma unroll
        for (int op = 0; op < 2; ++op)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                hv8 hi, lo;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float x = op ? vb[c][i] : va[c][i], sc = op ? scB : scA;
                    const _Float16 h = (_Float16)(x * sc);
                    hi[i] = h;
                    lo[i] = (_Float16)fmaf(x, sc, -(float)h);
                }
                *reinterpret_cast<hv8 *>(&Ls[buf][op][c][0][se * HST + 8 * sg]) = hi;
                *reinterpret_cast<hv8 *>(&Ls[buf][op][c][1][se * HST + 8 * sg]) = lo;
            }
    };

    // accumulators: Ar Br, Ai Bi (subtracted at the end), Ar Bi + Ai Br
    fv4 acp[7], acn[7], aci[7];
#pragma unroll
    for (int j = 0; j < 7; ++j)
        for (int q = 0; q < 4; ++q) { acp[j][q] = 0.0f; acn[j][q] = 0.0f; aci[j][q] = 0.0f; }

    const unsigned nstage = (ke > kb) ? (ke - kb + HK - 1) / HK : 0u;
    if (nstage > 0) {
        gload(kb);
        split_store(0);
        gload(kb + HK);   // past the end: reads 0
        __syncthreads();
    }
    for (unsigned s = 0; s < nstage; ++s) {
        const int buf = (int)(s & 1u);
        // the next stage's split and the loads two stages ahead run
        // unconditionally, so that the loop body is one basic block with the
        // MFMAs; past the last stage they read 0 and write the buffer that is
        // no longer read
        split_store(buf ^ 1);
        gload(kb + (s + 2) * HK);
        const int fo = (16 * wave + (lane & 15)) * HST + 8 * (lane >> 4);
        const hv8 arh = *reinterpret_cast<const hv8 *>(&Ls[buf][0][0][0][fo]);
        const hv8 arl = *reinterpret_cast<const hv8 *>(&Ls[buf][0][0][1][fo]);
        const hv8 aih = *reinterpret_cast<const hv8 *>(&Ls[buf][0][1][0][fo]);
        const hv8 ail = *reinterpret_cast<const hv8 *>(&Ls[buf][0][1][1][fo]);
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int bo = (16 * j + (lane & 15)) * HST + 8 * (lane >> 4);
            const hv8 brh = *reinterpret_cast<const hv8 *>(&Ls[buf][1][0][0][bo]);
            const hv8 brl = *reinterpret_cast<const hv8 *>(&Ls[buf][1][0][1][bo]);
            const hv8 bih = *reinterpret_cast<const hv8 *>(&Ls[buf][1][1][0][bo]);
            const hv8 bil = *reinterpret_cast<const hv8 *>(&Ls[buf][1][1][1][bo]);
            acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brh, acp[j], 0, 0, 0);
            acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bih, acn[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bih, aci[j], 0, 0, 0);
            acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brl, acp[j], 0, 0, 0);
            acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bil, acn[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bil, aci[j], 0, 0, 0);
            acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, brh, acp[j], 0, 0, 0);
            acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, bih, acn[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, bih, aci[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, brh, aci[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, brl, aci[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, brh, aci[j], 0, 0, 0);
        }
        __syncthreads();   // next buffer complete, this one free
    }
    const double back = ldexp(1.0, -(eA + eB));
    double *pr_ = part + (((size_t)split * nA2 + ij) * 2 + 0) * Mpad * Mpad;
    double *pi_ = pr_ + (size_t)Mpad * Mpad;
#pragma unroll
    for (int j = 0; j < 7; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + wave * 16 + 4 * (lane >> 4) + q;
            const int col = c0 + j * 16 + (lane & 15);
            pr_[(size_t)row * Mpad + col] = ((double)acp[j][q] - (double)acn[j][q]) * back;
            pi_[(size_t)row * Mpad + col] = (double)aci[j][q] * back;
        }
}

// The same contraction for panels that the Green function kernels already
// split (rows of [nA2][re|im][Mpad] words holding the fp16 pair (hi, lo) of
// x 2^e = hi + lo, bitwise the values contract_h2_kernel forms from the FP32
// panels).  A stage arrives 16 bytes (4 entries) per load and is separated
// into hi and lo with byte permutes: LDS holds [buffer][operand][re|im]
// [hi|lo][32 pairs][128 entries] fp16 (131,072 B), and the MFMA operands,
// which want 8 consecutive pairs of one entry per lane, come from transposing
// reads (ds_read_b64_tr_b16: within each 16 lanes, lane i receives element
// i & 3 of the 8-byte loads of lanes 4 j + (i >> 2), j = 0..3; lane t of a
// group addresses pair 4 h + (t >> 2), entries 4 (t & 3) .. +3, so lane i
// gets entry i of pairs 4 h .. 4 h + 3).  The 8-byte chunks of pair p's row
// sit XOR-swizzled, chunk c at c ^ (4 (p & 3) + 16 ((p >> 3) & 1)), so that
// the 32 lanes of two groups read 32 distinct chunk positions (no bank
// conflicts with 32 or 64 banks).  No conversion work remains in the loop.
typedef short sv4 __attribute__((ext_vector_type(4)));
typedef _Float16 hv4 __attribute__((ext_vector_type(4)));

constexpr int HROW = 128;   // LDS entries per pair row (112 used, 32 swizzled chunks of 4)

__device__ inline hv8 tr_frag(const _Float16 *base)
{
    // base: this lane's address for pairs 4 h + (t >> 2) of the first half;
    // the second half is 4 pairs on (same swizzle: p & 3 and p >> 3 agree)
    typedef __attribute__((address_space(3))) sv4 lsv4;
    const sv4 a = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lsv4 *)(base));
    const sv4 b = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lsv4 *)(base + 4 * HROW));
    const hv4 ha = __builtin_bit_cast(hv4, a), hb = __builtin_bit_cast(hv4, b);
    return hv8{ha[0], ha[1], ha[2], ha[3], hb[0], hb[1], hb[2], hb[3]};
}

__global__ __launch_bounds__(NTH) void contract_h2p_kernel(const unsigned *__restrict__ A,
                                                           const unsigned *__restrict__ B, long np, int Mpad,
                                                           int nA2, int ntile, int splits, long kper, int eA, int eB,
                                                           const double *__restrict__ p2, int n_x, double areal,
                                                           double *__restrict__ part)
{
    // [buffer][operand][re|im][hi|lo][pair * 128 + swizzled entry]
    __shared__ __align__(16) _Float16 Ls[2][2][2][2][HK * HROW];
    const int nwg = gridDim.x;
    const int b = blockIdx.x;
    const int n8 = nwg & ~7;
    const int lid = (b < n8) ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
    const int tile = lid % (ntile * ntile);
    const int rest = lid / (ntile * ntile);
    const int split = rest % splits;
    const int ij = rest / splits;
    const int tm = tile / ntile, tn = tile - tm * ntile;
    const int r0 = tm * TH_TILE, c0 = tn * TH_TILE;
    const unsigned kb = (unsigned)(split * kper);
    const unsigned ke = (unsigned)min(np, (long)kb + kper);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    auto swz = [](int p) { return 4 * (p & 3) + 16 * ((p >> 3) & 1); };
    // raw buffer loads over the split's rows (an offset at or past row ke
    // reads 0); the B resource starts (c0 - r0) words off B so that one
    // offset serves both operands.  Panel bytes + 96 rows < 2^31 (host check).
    const unsigned ps2 = (unsigned)nA2 * 2u * (unsigned)Mpad * 4u;   // bytes per pair row
    const int nrec = (int)(ke * ps2);
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc((void *)A, (short)0, nrec, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(reinterpret_cast<const char *>(B) + (long)(c0 - r0) * 4), (short)0, nrec, 0x00020000);
    if (p2) eA = th_real_exp(p2, n_x, areal);

    // staging: per operand a stage is 32 pairs x 2 components x 28 chunks of
    // 16 bytes (4 entries); thread tid copies chunks u * 448 + tid, u < 4
    unsigned goff[4], loff[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int q = u * NTH + tid;
        const int pr = q / 56, rem = q - pr * 56, c = rem / 28, ch = rem - c * 28;
        goff[u] = (unsigned)pr * ps2 + (((unsigned)ij * 2u + (unsigned)c) * (unsigned)Mpad + (unsigned)(r0 + 4 * ch)) * 4u;
        loff[u] = (unsigned)c * (2 * HK * HROW) + (unsigned)(pr * HROW + 4 * (ch ^ swz(pr)));
    }
    uint4 sa[4], sb[4];
    auto gload = [&](unsigned k0) {
        const unsigned st = k0 * ps2;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sa[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rA, (int)(goff[u] + st), 0, 0));
            sb[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rB, (int)(goff[u] + st), 0, 0));
        }
    };
    // words (hi | lo << 16) of 4 entries -> hi[4] and lo[4]
    auto put = [&](_Float16 *hl, const uint4 v) {
        const uint2 hi = make_uint2(__builtin_amdgcn_perm(v.y, v.x, 0x05040100u), __builtin_amdgcn_perm(v.w, v.z, 0x05040100u));
        const uint2 lo = make_uint2(__builtin_amdgcn_perm(v.y, v.x, 0x07060302u), __builtin_amdgcn_perm(v.w, v.z, 0x07060302u));
        *reinterpret_cast<uint2 *>(hl) = hi;
        *reinterpret_cast<uint2 *>(hl + HK * HROW) = lo;
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            put(&Ls[buf][0][0][0][loff[u]], sa[u]);
            put(&Ls[buf][1][0][0][loff[u]], sb[u]);
        }
    };

    fv4 acp[7], acn[7], aci[7];
#pragma unroll
    for (int j = 0; j < 7; ++j)
        for (int q = 0; q < 4; ++q) { acp[j][q] = 0.0f; acn[j][q] = 0.0f; aci[j][q] = 0.0f; }

    // this lane's transposing-read offsets within an array: pair
    // 8 (lane >> 4) + ((lane & 15) >> 2), chunk 4 b + (lane & 3) of the
    // 16-entry block b, swizzled: (4 b + k) ^ s = (4 b ^ s) + k (s has no
    // bits below 2)
    const int trp = (8 * (lane >> 4) + ((lane & 15) >> 2)) * HROW, trk = lane & 3;
    const int trs = swz(8 * (lane >> 4) + ((lane & 15) >> 2));
    auto tro = [&](int blk) { return trp + 4 * (((4 * blk) ^ trs) + trk); };
    const unsigned nstage = (ke > kb) ? (ke - kb + HK - 1) / HK : 0u;
    if (nstage > 0) {
        gload(kb);
        sstore(0);
        gload(kb + HK);   // past the end: reads 0
        __syncthreads();
    }
    for (unsigned s = 0; s < nstage; ++s) {
        const int buf = (int)(s & 1u);
        // the next stage's copy and the loads two stages ahead run
        // unconditionally (one basic block with the MFMAs); past the last
        // stage they read 0 and write the buffer that is no longer read
        sstore(buf ^ 1);
        gload(kb + (s + 2) * HK);
        // keep the loads here, ahead of the MFMAs: they then have the whole
        // stage to arrive before the next copy waits for them
        __builtin_amdgcn_sched_barrier(0);
        const int fo = tro(wave);
        const hv8 arh = tr_frag(&Ls[buf][0][0][0][fo]);
        const hv8 arl = tr_frag(&Ls[buf][0][0][1][fo]);
        const hv8 aih = tr_frag(&Ls[buf][0][1][0][fo]);
        const hv8 ail = tr_frag(&Ls[buf][0][1][1][fo]);
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int bo = tro(j);
            const hv8 brh = tr_frag(&Ls[buf][1][0][0][bo]);
            const hv8 brl = tr_frag(&Ls[buf][1][0][1][bo]);
            const hv8 bih = tr_frag(&Ls[buf][1][1][0][bo]);
            const hv8 bil = tr_frag(&Ls[buf][1][1][1][bo]);
            acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brh, acp[j], 0, 0, 0);
            acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bih, acn[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bih, aci[j], 0, 0, 0);
            acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brl, acp[j], 0, 0, 0);
            acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bil, acn[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bil, aci[j], 0, 0, 0);
            acp[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, brh, acp[j], 0, 0, 0);
            acn[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, bih, acn[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, bih, aci[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, brh, aci[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, brl, aci[j], 0, 0, 0);
            aci[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, brh, aci[j], 0, 0, 0);
        }
        __syncthreads();   // next buffer complete, this one free
    }
    const double back = ldexp(1.0, -(eA + eB));
    double *pr_ = part + (((size_t)split * nA2 + ij) * 2 + 0) * Mpad * Mpad;
    double *pi_ = pr_ + (size_t)Mpad * Mpad;
#pragma unroll
    for (int j = 0; j < 7; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + wave * 16 + 4 * (lane >> 4) + q;
            const int col = c0 + j * 16 + (lane & 15);
            pr_[(size_t)row * Mpad + col] = ((double)acp[j][q] - (double)acn[j][q]) * back;
            pi_[(size_t)row * Mpad + col] = (double)aci[j][q] * back;
        }
}

// Building blocks of the eight-wave contraction below: one 16 x 16 complex
// block's operands (re/im x hi/lo fragments, transposing reads) and its 12
// fp16 MFMAs (hi.hi + hi.lo + lo.hi per real product, Ar Br and Ai Bi in
// separate accumulators).
struct h2_frags { hv8 rh, rl, ih, il; };

__device__ __forceinline__ h2_frags h2_load(const _Float16 *L, int o)
{
    // L: [re|im][hi|lo][HK * HROW] of one operand in one buffer
    h2_frags f;
    f.rh = tr_frag(L + o);
    f.rl = tr_frag(L + HK * HROW + o);
    f.ih = tr_frag(L + 2 * HK * HROW + o);
    f.il = tr_frag(L + 3 * HK * HROW + o);
    return f;
}

__device__ __forceinline__ void h2_mac(const h2_frags &a, const h2_frags &b, fv4 &p, fv4 &n, fv4 &i)
{
    p = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.rh, b.rh, p, 0, 0, 0);
    n = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.ih, b.ih, n, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.rh, b.ih, i, 0, 0, 0);
    p = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.rh, b.rl, p, 0, 0, 0);
    n = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.ih, b.il, n, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.rh, b.il, i, 0, 0, 0);
    p = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.rl, b.rh, p, 0, 0, 0);
    n = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.il, b.ih, n, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.rl, b.ih, i, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.ih, b.rh, i, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.ih, b.rl, i, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.il, b.rh, i, 0, 0, 0);
}

// one wave's rectangle of blocks: block rows R0 .. R0 + NR - 1, block
// columns C0 .. C0 + NC - 1 of the tile
template <int R0, int NR, int C0, int NC> struct h2_rect {
    fv4 p[NR][NC], n[NR][NC], i[NR][NC];
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) { p[r][c][q] = 0.0f; n[r][c][q] = 0.0f; i[r][c][q] = 0.0f; }
    }
    template <typename TRO>
    __device__ __forceinline__ void step(const _Float16 *LA, const _Float16 *LB, TRO tro)
    {
        h2_frags a[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) a[r] = h2_load(LA, tro(R0 + r));
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const h2_frags b = h2_load(LB, tro(C0 + c));
#pragma unroll
            for (int r = 0; r < NR; ++r) h2_mac(a[r], b, p[r][c], n[r][c], i[r][c]);
        }
    }
    __device__ __forceinline__ void store(double *pr_, double *pi_, int Mpad, int r0, int c0, int lane, double back)
    {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = r0 + (R0 + r) * 16 + 4 * (lane >> 4) + q;
                    const int col = c0 + (C0 + c) * 16 + (lane & 15);
                    pr_[(size_t)row * Mpad + col] = ((double)p[r][c][q] - (double)n[r][c][q]) * back;
                    pi_[(size_t)row * Mpad + col] = (double)i[r][c][q] * back;
                }
    }
};

// The same contraction with eight waves, two per SIMD (one wave's waits
// are covered by the other's MFMAs), each holding a small 2-D block of the
// 7 x 7 tile:
//   waves 0-5: block rows 2 (w mod 3) .. +1 x block columns 3 (w / 3) .. +2  (6 blocks)
//   wave 6:    block rows 0-5 x block column 6                              (6)
//   wave 7:    block row 6 x block columns 0-6                              (7)
// SIMD s runs waves s and s + 4: 12, 12, 12 and 13 blocks.  Per stage the
// waves read 45 operand blocks from LDS (the seven-strip layout: 56).  A
// stage is 1792 chunks of 16 bytes, 3.5 per thread (the fourth copy is
// done by the first 256 threads).
constexpr int NTH8 = 512;

template <int W>
__device__ __forceinline__ void h2e_body(const unsigned *__restrict__ A, const unsigned *__restrict__ B, long np,
                                         int Mpad, int nA2, int ntile, int splits, long kper, int eA, int eB,
                                         double *__restrict__ part, _Float16 (*Ls)[2][2][2][HK * HROW])
{
    const int nwg = gridDim.x;
    const int b = blockIdx.x;
    const int n8 = nwg & ~7;
    const int lid = (b < n8) ? (b & 7) * (n8 >> 3) + (b >> 3) : b;
    const int tile = lid % (ntile * ntile);
    const int rest = lid / (ntile * ntile);
    const int split = rest % splits;
    const int ij = rest / splits;
    const int tm = tile / ntile, tn = tile - tm * ntile;
    const int r0 = tm * TH_TILE, c0 = tn * TH_TILE;
    const unsigned kb = (unsigned)(split * kper);
    const unsigned ke = (unsigned)min(np, (long)kb + kper);
    const int tid = threadIdx.x, lane = tid & 63;
    auto swz = [](int p) { return 4 * (p & 3) + 16 * ((p >> 3) & 1); };
    const unsigned ps2 = (unsigned)nA2 * 2u * (unsigned)Mpad * 4u;   // bytes per pair row
    const int nrec = (int)(ke * ps2);
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc((void *)A, (short)0, nrec, 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(reinterpret_cast<const char *>(B) + (long)(c0 - r0) * 4), (short)0, nrec, 0x00020000);

    // staging: chunk u * 512 + tid of the stage's 1792 (u < 3 for every
    // thread, u = 3 for tid < 256)
    constexpr int NU = 4;
    const bool last = tid < 1792 - 3 * NTH8;
    unsigned goff[NU], loff[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int q = u < 3 || last ? u * NTH8 + tid : 0;
        const int pr = q / 56, rem = q - pr * 56, c = rem / 28, ch = rem - c * 28;
        goff[u] = (unsigned)pr * ps2 + (((unsigned)ij * 2u + (unsigned)c) * (unsigned)Mpad + (unsigned)(r0 + 4 * ch)) * 4u;
        loff[u] = (unsigned)c * (2 * HK * HROW) + (unsigned)(pr * HROW + 4 * (ch ^ swz(pr)));
    }
    // two register sets: stage t is loaded into set t & 1 three stages
    // before it is computed and copied to LDS one stage before (global loads
    // under this traffic take longer than one stage of MFMA work)
    uint4 sa2[2][NU], sb2[2][NU];
    auto gload = [&](int set, unsigned k0) {
        uint4(&sa)[NU] = sa2[set];
        uint4(&sb)[NU] = sb2[set];
        const unsigned st = k0 * ps2;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            sa[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rA, (int)(goff[u] + st), 0, 0));
            sb[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rB, (int)(goff[u] + st), 0, 0));
        }
        if (last) {
            sa[3] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rA, (int)(goff[3] + st), 0, 0));
            sb[3] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rB, (int)(goff[3] + st), 0, 0));
        }
    };
    auto put = [&](_Float16 *hl, const uint4 v) {
        const uint2 hi = make_uint2(__builtin_amdgcn_perm(v.y, v.x, 0x05040100u), __builtin_amdgcn_perm(v.w, v.z, 0x05040100u));
        const uint2 lo = make_uint2(__builtin_amdgcn_perm(v.y, v.x, 0x07060302u), __builtin_amdgcn_perm(v.w, v.z, 0x07060302u));
        *reinterpret_cast<uint2 *>(hl) = hi;
        *reinterpret_cast<uint2 *>(hl + HK * HROW) = lo;
    };
    auto sstore = [&](int buf, int set) {
        const uint4(&sa)[NU] = sa2[set];
        const uint4(&sb)[NU] = sb2[set];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            put(&Ls[buf][0][0][0][loff[u]], sa[u]);
            put(&Ls[buf][1][0][0][loff[u]], sb[u]);
        }
        if (last) {
            put(&Ls[buf][0][0][0][loff[3]], sa[3]);
            put(&Ls[buf][1][0][0][loff[3]], sb[3]);
        }
    };
    const int trp = (8 * (lane >> 4) + ((lane & 15) >> 2)) * HROW, trk = lane & 3;
    const int trs = swz(8 * (lane >> 4) + ((lane & 15) >> 2));
    auto tro = [=](int blk) { return trp + 4 * (((4 * blk) ^ trs) + trk); };

    // this wave's blocks
    constexpr int R0 = W < 6 ? 2 * (W % 3) : W == 6 ? 0 : 6;
    constexpr int NR = W < 6 ? 2 : W == 6 ? 6 : 1;
    constexpr int C0 = W < 6 ? 3 * (W / 3) : W == 6 ? 6 : 0;
    constexpr int NC = W < 6 ? 3 : W == 6 ? 1 : 7;
    h2_rect<R0, NR, C0, NC> m;
    m.zero();

    const unsigned nstage = (ke > kb) ? (ke - kb + HK - 1) / HK : 0u;
    auto stage = [&](unsigned s, int odd) {
        // stage s + 1 (set odd ^ 1) to LDS, stage s + 3 into that set; past
        // the last stage the loads read 0 and the copy fills a buffer no
        // longer read
        sstore(odd ^ 1, odd ^ 1);
        gload(odd ^ 1, kb + (s + 3) * HK);
        __builtin_amdgcn_sched_barrier(0);
        m.step(&Ls[odd][0][0][0][0], &Ls[odd][1][0][0][0], tro);
        __syncthreads();   // next buffer complete, this one free
    };
    if (nstage > 0) {
        gload(0, kb);
        sstore(0, 0);
        gload(1, kb + HK);
        gload(0, kb + 2 * HK);
        __syncthreads();
    }
    // unrolled by two so that the register set of every copy is known at
    // compile time
    unsigned s = 0;
    for (; s + 1 < nstage; s += 2) {
        stage(s, 0);
        stage(s + 1, 1);
    }
    if (s < nstage) stage(s, 0);
    const double back = ldexp(1.0, -(eA + eB));
    double *pr_ = part + (((size_t)split * nA2 + ij) * 2 + 0) * Mpad * Mpad;
    double *pi_ = pr_ + (size_t)Mpad * Mpad;
    m.store(pr_, pi_, Mpad, r0, c0, lane, back);
}

__global__ __launch_bounds__(NTH8) __attribute__((amdgpu_waves_per_eu(2, 2))) void contract_h2e_kernel(
    const unsigned *__restrict__ A, const unsigned *__restrict__ B, long np, int Mpad, int nA2, int ntile, int splits,
    long kper, int eA, int eB, const double *__restrict__ p2, int n_x, double areal, double *__restrict__ part)
{
    __shared__ __align__(16) _Float16 Ls[2][2][2][2][HK * HROW];
    if (p2) eA = th_real_exp(p2, n_x, areal);
    switch (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) {
    case 0: h2e_body<0>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    case 1: h2e_body<1>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    case 2: h2e_body<2>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    case 3: h2e_body<3>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    case 4: h2e_body<4>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    case 5: h2e_body<5>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    case 6: h2e_body<6>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    default: h2e_body<7>(A, B, np, Mpad, nA2, ntile, splits, kper, eA, eB, part, Ls); break;
    }
}

// Adjoint of the contraction (reverse mode).  With a cotangent That of one
// axis's accumulator, T = sum_p A_p B_p^T gives
//     Ahat_p = That conj(B_p),   Bhat_p = That^T conj(A_p)
// per atom pair: a GEMM whose rows are the pairs of the batch and whose
// inner dimension is the flattened block (K = Mpad).  out[p][n] =
// sum_k conj(X[p][k]) Y[k][n] with Y[k][n] = That[n][k] (trans = 1, for
// Ahat from X = B) or That[k][n] (trans = 0, for Bhat from X = A).  A
// workgroup owns 112 pairs x 112 columns of one atom pair; the seven waves
// each hold 16 pairs as 7 MFMA blocks, the same register layout as the
// forward kernel.  In the mixed mode That is rounded to FP32 in LDS.
template <typename P, bool K3M>
__global__ __launch_bounds__(NTH) void panel_adjoint_kernel(const P *__restrict__ X, const double *__restrict__ That,
                                                            long np, int Mpad, int nA2, int ntile, int trans,
                                                            double *__restrict__ out, th_rowmap rows,
                                                            float unsplit)
{
    __shared__ __align__(16) P Xs[2][TH_KC][2][TH_TILE];
    __shared__ __align__(16) P Ys[2][TH_KC][2][TH_TILE];
    constexpr int STAGE = TH_KC * 2 * TH_TILE;            // elements per matrix per stage
    constexpr int PER_T = (STAGE + NTH - 1) / NTH;
    const long nptile = (np + TH_TILE - 1) / TH_TILE;
    const long b = blockIdx.x;
    const int tn = (int)(b % ntile);
    const long rest = b / ntile;
    const long tp = rest % nptile;
    const int ij = (int)(rest / nptile);
    const long p0 = tp * TH_TILE;
    const int n0 = tn * TH_TILE;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const size_t pstride = (size_t)nA2 * 2 * Mpad;
    const size_t MM = (size_t)Mpad * Mpad;
    const P *Xb = X + (size_t)ij * 2 * Mpad;
    const double *Tb = That + (size_t)ij * 2 * MM;

    typename mfma_t<P>::acc accr[7], acci[7], acc3[K3M ? 7 : 1];
    for (int j = 0; j < 7; ++j)
        for (int q = 0; q < 4; ++q) { accr[j][q] = 0; acci[j][q] = 0; }
    for (int j = 0; j < (K3M ? 7 : 1); ++j)
        for (int q = 0; q < 4; ++q) acc3[j][q] = 0;

    P rx[PER_T], ry[PER_T];
    long xrow[PER_T];   // panel row of each staged pair (fixed over the stages)
    // Stage element v of X: kk fastest (8 consecutive k of one pair row are
    // one contiguous segment of the panel), then the 112 rows, then re/im.
    // Y follows its memory order: kk fastest for trans = 1 (rows n of That),
    // n fastest for trans = 0 (rows k of That).
    auto xidx = [](int v, int &kk, int &m, int &comp) {
        kk = v % TH_KC; m = (v / TH_KC) % TH_TILE; comp = v / (TH_KC * TH_TILE);
    };
    auto yidx = [trans](int v, int &kk, int &m, int &comp) {
        comp = v / (TH_KC * TH_TILE);
        const int r = v - comp * TH_KC * TH_TILE;
        if (trans) { kk = r % TH_KC; m = r / TH_KC; }
        else { m = r % TH_TILE; kk = r / TH_TILE; }
    };
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            const int v = tid + u * NTH;
            P x = 0, y = 0;
            if (v < STAGE) {
                int kk, m, comp;
                xidx(v, kk, m, comp);
                if (p0 + m < np) x = Xb[(size_t)xrow[u] * pstride + (size_t)comp * Mpad + k0 + kk];
                if constexpr (sizeof(P) == sizeof(float)) {
                    // split panels: the word holds fp16 (hi, lo) of x 2^e; unsplit = 2^-e
                    if (unsplit != 0.0f) {
                        const unsigned wv = __builtin_bit_cast(unsigned, (float)x);
                        const float hi = (float)__builtin_bit_cast(_Float16, (unsigned short)(wv & 0xffffu));
                        const float lo = (float)__builtin_bit_cast(_Float16, (unsigned short)(wv >> 16));
                        x = (P)((hi + lo) * unsplit);
                    }
                }
                yidx(v, kk, m, comp);
                const int n = n0 + m, k = k0 + kk;
                y = (P)(trans ? Tb[(size_t)comp * MM + (size_t)n * Mpad + k] : Tb[(size_t)comp * MM + (size_t)k * Mpad + n]);
            }
            rx[u] = x;
            ry[u] = y;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < PER_T; ++u) {
            const int v = tid + u * NTH;
            if (v < STAGE) {
                int kk, m, comp;
                xidx(v, kk, m, comp);
                Xs[buf][kk][comp][m] = rx[u];
                yidx(v, kk, m, comp);
                Ys[buf][kk][comp][m] = ry[u];
            }
        }
    };
#pragma unroll
    for (int u = 0; u < PER_T; ++u) {
        int kk, m, comp;
        xidx(tid + u * NTH, kk, m, comp);
        xrow[u] = th_map_row(rows.pg0 + p0 + m, rows);
    }
    const int nstage = Mpad / TH_KC;
    gload(0);
    for (int s = 0; s < nstage; ++s) {
        const int buf = s & 1;
        sstore(buf);
        __syncthreads();
        if (s + 1 < nstage) gload((s + 1) * TH_KC);
#pragma unroll
        for (int kk = 0; kk < TH_KC; kk += 4) {
            const int pr = kk + (lane >> 4);
            const int m = wave * 16 + (lane & 15);
            const P xr = Xs[buf][pr][0][m], xi = Xs[buf][pr][1][m];
            if constexpr (K3M) {
                // P1 = xr yr, P2 = xi yi, P3 = (xr - xi)(yr + yi):
                // conj(x) y = (P1 + P2) + i (P3 - P1 + P2)
                const P xd = xr - xi;
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const int n = j * 16 + (lane & 15);
                    const P yr = Ys[buf][pr][0][n], yi = Ys[buf][pr][1][n];
                    accr[j] = mfma_t<P>::mac(xr, yr, accr[j]);
                    acci[j] = mfma_t<P>::mac(xi, yi, acci[j]);
                    acc3[j] = mfma_t<P>::mac(xd, yr + yi, acc3[j]);
                }
            } else {
                const P nxi = -xi;
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const int n = j * 16 + (lane & 15);
                    const P yr = Ys[buf][pr][0][n], yi = Ys[buf][pr][1][n];
                    // conj(x) * y = (xr yr + xi yi) + i (xr yi - xi yr)
                    accr[j] = mfma_t<P>::mac(xr, yr, accr[j]);
                    accr[j] = mfma_t<P>::mac(xi, yi, accr[j]);
                    acci[j] = mfma_t<P>::mac(xr, yi, acci[j]);
                    acci[j] = mfma_t<P>::mac(nxi, yr, acci[j]);
                }
            }
        }
        // the next iteration's sstore overwrites the other buffer only
    }
#pragma unroll
    for (int j = 0; j < 7; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long pp = p0 + wave * 16 + mfma_t<P>::row(lane >> 4, q);
            const int col = n0 + j * 16 + (lane & 15);
            if (pp < np) {
                double *o = out + (size_t)pp * pstride + (size_t)ij * 2 * Mpad;
                if constexpr (K3M) {
                    const double p1 = (double)accr[j][q], p2 = (double)acci[j][q], p3 = (double)acc3[j][q];
                    o[col] = p1 + p2;
                    o[Mpad + col] = p3 - p1 + p2;
                } else {
                    o[col] = (double)accr[j][q];
                    o[Mpad + col] = (double)acci[j][q];
                }
            }
        }
}

// Largest |x| of a buffer as the bits of a non-negative float (atomicMax on
// the bits: they order as the values); split words count as (hi + lo) u.
// The workgroup's largest value, one atomicMax per workgroup (atomics of
// every wave on the one address serialise and would dominate the scan).
__device__ inline void block_max_to(float m, unsigned *mx)
{
    __shared__ float wm[4];
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, wm[w]);
        atomicMax(mx, __builtin_bit_cast(unsigned, m));
    }
}

__global__ __launch_bounds__(256) void absmax_f_kernel(const float *__restrict__ x, size_t n, float u,
                                                       unsigned *__restrict__ mx)
{
    float m = 0.0f;
    auto val = [u](float v) {
        if (u != 0.0f) {
            const unsigned w = __builtin_bit_cast(unsigned, v);
            v = ((float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu)) +
                 (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16))) * u;
        }
        return fabsf(v);
    };
    auto vmax = [&](float4 v) { return fmaxf(fmaxf(val(v.x), val(v.y)), fmaxf(val(v.z), val(v.w))); };
    // 16-byte loads (n is a multiple of 4: whole panel rows of 2 Mpad floats),
    // four of them in flight per thread
    const float4 *x4 = reinterpret_cast<const float4 *>(x);
    const size_t n4 = n / 4, st = (size_t)gridDim.x * blockDim.x;
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; e + 3 * st < n4; e += 4 * st) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = x4[e + u * st];
#pragma unroll
        for (int u = 0; u < 4; ++u) m = fmaxf(m, vmax(v[u]));
    }
    for (; e < n4; e += st) m = fmaxf(m, vmax(x4[e]));
    block_max_to(m, mx);
}

// mx[0] = the bits of the largest of n non-negative floats (one workgroup)
__global__ __launch_bounds__(256) void max_of_kernel(const float *__restrict__ x, long n, unsigned *__restrict__ mx)
{
    float m = 0.0f;
    for (long e = threadIdx.x; e < n; e += blockDim.x) m = fmaxf(m, x[e]);
    block_max_to(m, mx);
}

__global__ __launch_bounds__(256) void absmax_d_kernel(const double *__restrict__ x, size_t n, unsigned *__restrict__ mx)
{
    float m = 0.0f;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, (float)fabs(x[e]));
    block_max_to(m, mx);
}

__global__ void zero_u_kernel(unsigned *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// power of two 2^e with |x| 2^e < 2^14 for |x| <= the maximum held in bits
__device__ inline int split_exp(unsigned bits)
{
    const float m = __builtin_bit_cast(float, bits);
    if (!(m > 0.0f) || !isfinite(m)) return 0;
    return min(100, max(-100, 13 - ilogbf(m)));
}

// That of one axis split into fp16 planes in the layout the fp16 panel
// adjoint's B operand reads, Ys[ij][re|im][hi|lo][n][k] with k padded to Kp
// (zeros), Y[k][n] = That[n][k] (trans = 1) or That[k][n] (trans = 0), under
// the power of two of the largest |That| (mx[1])
__global__ __launch_bounds__(256) void presplit_y_kernel(const double *__restrict__ That, int Mpad, int Kp, int nA2,
                                                         int trans, const unsigned *__restrict__ mx,
                                                         _Float16 *__restrict__ Ys)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nk = (size_t)Mpad * Kp;
    if (e >= (size_t)nA2 * 2 * nk) return;
    const size_t blk = e / nk;                  // ij * 2 + comp
    const int nkk = (int)(e - blk * nk), n = nkk / Kp, k = nkk - n * Kp;
    const size_t MM = (size_t)Mpad * Mpad;
    float v = 0.0f;
    if (k < Mpad) {
        const double y = trans ? That[blk * MM + (size_t)n * Mpad + k] : That[blk * MM + (size_t)k * Mpad + n];
        v = (float)(y * ldexp(1.0, split_exp(mx[1])));
    }
    const _Float16 hi = (_Float16)v;
    _Float16 *o = Ys + blk * 2 * nk + nkk;
    o[0] = hi;
    o[nk] = (_Float16)fmaf(v, 1.0f, -(float)hi);
}

// The adjoint's GEMM (out[p][n] = sum_k conj(X[p][k]) Y[k][n], as
// panel_adjoint_kernel) on the fp16 matrix cores for the mixed path: X (the
// batch's FP32 panels, or split words read as (hi + lo) u) and Y (That in
// FP64) are split into fp16 hi + lo under power-of-two scales from their
// largest entries (mx[0], mx[1], set by the absmax kernels; split words keep
// their exponent, ex_fixed), conj(X) by its
// negated imaginary part, and each real product is hi.hi + hi.lo + lo.hi in
// FP32 accumulators (Ar Br and Ai Bi apart).  A workgroup owns 112 pairs x 112
// columns of one atom pair, seven waves of 16 pairs x 7 column blocks; K (the
// panel index) runs in stages of 32 through LDS, each thread splitting groups
// of 8 consecutive k into one 16-byte LDS store per part; loads use clamped
// indices (no branch per load) and zero what lies outside.
constexpr int PAK = 32, PAST = PAK + 8;   // K per stage, LDS row stride in halves
__global__ __launch_bounds__(448) __attribute__((amdgpu_waves_per_eu(4))) void panel_adjoint_h2_kernel(const float *__restrict__ X,
                                                               const _Float16 *__restrict__ Ys, int Kp, long np, int Mpad,
                                                               int nA2, int ntile, int trans, double *__restrict__ out,
                                                               th_rowmap rows, float unsplit, int ex_fixed,
                                                               const unsigned *__restrict__ mx)
{
    typedef _Float16 hv8_t __attribute__((ext_vector_type(8)));
    __shared__ __align__(16) _Float16 Lx[2][2][TH_TILE * PAST];   // [re|im][hi|lo][pair][k]
    __shared__ __align__(16) _Float16 Ly[2][2][TH_TILE * PAST];   // [re|im][hi|lo][n][k]
    const long nptile = (np + TH_TILE - 1) / TH_TILE;
    const long b = blockIdx.x;
    const int tn = (int)(b % ntile);
    const long rest = b / ntile;
    const long tp = rest % nptile;
    const int ij = (int)(rest / nptile);
    const long p0 = tp * TH_TILE;
    const int n0 = tn * TH_TILE;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const size_t pstride = (size_t)nA2 * 2 * Mpad;
    // split words (unsplit != 0) keep their own exponent: (hi + lo) 2^e splits
    // back into the same two parts
    const int eX = ex_fixed != INT_MIN ? ex_fixed : split_exp(mx[0]), eY = split_exp(mx[1]);
    const float sx = ldexpf(1.0f, eX);
    const float *Xb = X + (size_t)ij * 2 * Mpad;
    const size_t nk = (size_t)Mpad * Kp;
    const _Float16 *Yb = Ys + (size_t)ij * 4 * nk;   // [re|im][hi|lo][n][k] of this atom pair
    // staging: 112 rows x 2 components x 4 groups of 8 k = 896 groups per
    // operand, two per thread
    auto split8 = [](const float (&v)[8], hv8_t &hi, hv8_t &lo) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const _Float16 h = (_Float16)v[i];
            hi[i] = h;
            lo[i] = (_Float16)fmaf(v[i], 1.0f, -(float)h);
        }
    };
    // real part: Ar Br and -Ai Bi in one accumulator, imaginary part Ar Bi + Ai Br
    fv4 ap[7], ai[7];
#pragma unroll
    for (int j = 0; j < 7; ++j)
        for (int q = 0; q < 4; ++q) { ap[j][q] = 0.0f; ai[j][q] = 0.0f; }
    long xrow[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int g = tid + 448 * u, m = g % TH_TILE;
        const long pp = p0 + m < np ? p0 + m : np - 1;
        xrow[u] = th_map_row(rows.pg0 + pp, rows);
    }
    // the next stage's X and Y in registers while the current one's products
    // run: X as two 16-byte loads per group of 8 k (Mpad is a multiple of 8, so
    // a group lies wholly inside or wholly past the panel row; past it the
    // load reads the row's last group and the values are zeroed)
    float4 xq[2][2];
    hv8_t yq[2][2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int g = tid + 448 * u, m = g % TH_TILE, r = g / TH_TILE, comp = r >> 2, kg = 8 * (r & 3);
            const int ks = k0 + kg < Mpad ? k0 + kg : Mpad - 8;
            const float4 *xp = reinterpret_cast<const float4 *>(Xb + (size_t)xrow[u] * pstride + (size_t)comp * Mpad + ks);
            xq[u][0] = xp[0];
            xq[u][1] = xp[1];
            const size_t yo = (size_t)comp * 2 * nk + (size_t)(n0 + m) * Kp + k0 + kg;
            yq[u][0] = *reinterpret_cast<const hv8_t *>(Yb + yo);
            yq[u][1] = *reinterpret_cast<const hv8_t *>(Yb + yo + nk);
        }
    };
    for (int k0 = 0; k0 < Mpad; k0 += PAK) {
        gload(k0);
        __syncthreads();                          // the previous stage's reads done
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int g = tid + 448 * u, m = g % TH_TILE, r = g / TH_TILE, comp = r >> 2, kg = 8 * (r & 3);
            const bool in = p0 + m < np && k0 + kg < Mpad;
            const float xs[8] = {xq[u][0].x, xq[u][0].y, xq[u][0].z, xq[u][0].w,
                                 xq[u][1].x, xq[u][1].y, xq[u][1].z, xq[u][1].w};
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                float x = xs[i];
                if (unsplit != 0.0f) {
                    const unsigned w = __builtin_bit_cast(unsigned, x);
                    x = ((float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu)) +
                         (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16))) * unsplit;
                }
                v[i] = in ? (comp ? -x : x) * sx : 0.0f;   // conj(X)
            }
            hv8_t hi, lo;
            split8(v, hi, lo);
            *reinterpret_cast<hv8_t *>(&Lx[comp][0][m * PAST + kg]) = hi;
            *reinterpret_cast<hv8_t *>(&Lx[comp][1][m * PAST + kg]) = lo;
            *reinterpret_cast<hv8_t *>(&Ly[comp][0][m * PAST + kg]) = yq[u][0];
            *reinterpret_cast<hv8_t *>(&Ly[comp][1][m * PAST + kg]) = yq[u][1];
        }
        __syncthreads();
        const int ao = (16 * wave + (lane & 15)) * PAST + 8 * (lane >> 4);
        const hv8_t arh = *reinterpret_cast<const hv8_t *>(&Lx[0][0][ao]);
        const hv8_t arl = *reinterpret_cast<const hv8_t *>(&Lx[0][1][ao]);
        const hv8_t aih = *reinterpret_cast<const hv8_t *>(&Lx[1][0][ao]);
        const hv8_t ail = *reinterpret_cast<const hv8_t *>(&Lx[1][1][ao]);
        const hv8_t nih = -aih, nil = -ail;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int bo = (16 * j + (lane & 15)) * PAST + 8 * (lane >> 4);
            const hv8_t brh = *reinterpret_cast<const hv8_t *>(&Ly[0][0][bo]);
            const hv8_t brl = *reinterpret_cast<const hv8_t *>(&Ly[0][1][bo]);
            const hv8_t bih = *reinterpret_cast<const hv8_t *>(&Ly[1][0][bo]);
            const hv8_t bil = *reinterpret_cast<const hv8_t *>(&Ly[1][1][bo]);
            ap[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brh, ap[j], 0, 0, 0);
            ap[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(nih, bih, ap[j], 0, 0, 0);
            ai[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bih, ai[j], 0, 0, 0);
            ap[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, brl, ap[j], 0, 0, 0);
            ap[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(nih, bil, ap[j], 0, 0, 0);
            ai[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bil, ai[j], 0, 0, 0);
            ap[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, brh, ap[j], 0, 0, 0);
            ap[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(nil, bih, ap[j], 0, 0, 0);
            ai[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, bih, ai[j], 0, 0, 0);
            ai[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, brh, ai[j], 0, 0, 0);
            ai[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, brl, ai[j], 0, 0, 0);
            ai[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, brh, ai[j], 0, 0, 0);
        }
    }
    const double back = ldexp(1.0, -(eX + eY));
#pragma unroll
    for (int j = 0; j < 7; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long pp = p0 + wave * 16 + 4 * (lane >> 4) + q;
            const int col = n0 + j * 16 + (lane & 15);
            if (pp < np) {
                double *o = out + (size_t)pp * pstride + (size_t)ij * 2 * Mpad;
                o[col] = (double)ap[j][q] * back;
                o[Mpad + col] = (double)ai[j][q] * back;
            }
        }
}


// acc += sum_s part[s], elementwise, fixed order
__global__ void reduce_kernel(const double *__restrict__ part, double *__restrict__ acc, int splits, size_t n)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[(size_t)k * n + e];
    acc[e] += s;
}

__global__ void zero_kernel(double *p, size_t n)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = 0.0;
}

// chi0[X,Y] with X = (sigma,i,mu,nu), Y = (sigma',j,gamma,xi) (column-major D x D)
__global__ void finalize_kernel(const double *__restrict__ acc, double2 *__restrict__ chi0, int nA, int nOrb,
                                int Mpad, int D)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)D * D) return;
    const int X = (int)(e % D), Y = (int)(e / D);
    const int bs = 2 * nOrb, nA2 = nA * nA;
    int s1 = X & 3, r = X >> 2;
    const int nu = r % nOrb; r /= nOrb;
    const int mu = r % nOrb;
    const int i = r / nOrb;
    int s2 = Y & 3;
    r = Y >> 2;
    const int xi = r % nOrb; r /= nOrb;
    const int ga = r % nOrb;
    const int j = r / nOrb;
    // spin labels: a = (dn,up,dn,up), c = (up,up,dn,dn) for sigma = (+,up,dn,-)
    const int a1 = (s1 == 0 || s1 == 2) ? 1 : 0, c1 = (s1 >= 2) ? 1 : 0;
    const int a2 = (s2 == 0 || s2 == 2) ? 1 : 0, c2 = (s2 >= 2) ? 1 : 0;
    const size_t MM = (size_t)Mpad * Mpad;
    const double *AI = acc;                              // imaginary axis [nA2][2][Mpad][Mpad]
    const double *AR = acc + (size_t)nA2 * 2 * MM;       // real axis
    // term 1: T(i,j)[r][c]
    const int rr = (a1 * nOrb + nu) * bs + (a2 * nOrb + ga);
    const int cc = (c2 * nOrb + xi) * bs + (c1 * nOrb + mu);
    const size_t b1 = (size_t)(i * nA + j) * 2 * MM + (size_t)rr * Mpad + cc;
    // term 2: conj(T(j,i)[r'][c'])
    const int rr2 = (a2 * nOrb + ga) * bs + (a1 * nOrb + nu);
    const int cc2 = (c1 * nOrb + mu) * bs + (c2 * nOrb + xi);
    const size_t b2 = (size_t)(j * nA + i) * 2 * MM + (size_t)rr2 * Mpad + cc2;
    double re = AI[b1] + AI[b2] + AR[b1];
    double im = AI[b1 + MM] - AI[b2 + MM] + AR[b1 + MM];
    const double inv2pi = 0.15915494309189533577;   // 1/(2 pi)
    chi0[e] = make_double2(re * inv2pi, im * inv2pi);
    (void)nA2;
}

}  // namespace

int th_contract_splits(const th_handle *h, long np, int prec)
{
    long s = h->splits;
    if (prec == TH_PREC_MIXED) {
        // measured on the 54 x 54 bcc supercell: the rounding of a split's
        // FP32 partial sums grows with its length (1.6e-7 relative at 7282
        // pairs, 8.8e-7 at 29127)
        const long lmax = 8192;
        s = std::max(s, (np + lmax - 1) / lmax);
    }
    return (int)std::min<long>(s, 4096);
}

namespace {

template <typename P>
int launch_contract_t(th_handle *h, const void *pA, const void *pB, long np, int axis, const th_contract_scale *sc)
{
    const int ntile = h->Mpad / TH_TILE;
    const int splits = th_contract_splits(h, np, sizeof(P) == sizeof(double) ? TH_PREC_FP64 : TH_PREC_MIXED);
    if (splits > h->part_cap) {
        th_set_error("contraction: partial buffer holds " + std::to_string(h->part_cap) + " splits, " +
                     std::to_string(splits) + " needed");
        return TH_E_STATE;
    }
    long kper = (np + splits - 1) / splits;
    kper = ((kper + TH_KC - 1) / TH_KC) * TH_KC;
    const int nwg = ntile * ntile * splits * h->nA2;
    // the split-panel kernels read up to 127 rows past a split's end (their
    // prefetch; such reads return 0): the buffer offsets stay below 2^31
    const bool fits = (unsigned long)(np + 128) * h->nA2 * 2 * h->Mpad * sizeof(float) < (1ul << 31);
    if (sc && sc->presplit) {
        if (sizeof(P) != sizeof(float) || !fits) {
            th_set_error("contraction: split half-precision panels need the mixed path and < 2 GiB per batch");
            return TH_E_STATE;
        }
        long kh = (np + splits - 1) / splits;
        kh = ((kh + HK - 1) / HK) * HK;
        if (h->contract_tiles)
            hipLaunchKernelGGL(contract_h2e_kernel, dim3(nwg), dim3(NTH8), 0, h->stream, (const unsigned *)pA,
                               (const unsigned *)pB, np, h->Mpad, h->nA2, ntile, splits, kh, sc->eA, sc->eB, sc->p2,
                               sc->n_x, sc->areal, h->d_part);
        else
            hipLaunchKernelGGL(contract_h2p_kernel, dim3(nwg), dim3(NTH), 0, h->stream, (const unsigned *)pA,
                               (const unsigned *)pB, np, h->Mpad, h->nA2, ntile, splits, kh, sc->eA, sc->eB, sc->p2,
                               sc->n_x, sc->areal, h->d_part);
    } else if (sizeof(P) == sizeof(float) && sc && h->contract_f16 && fits) {
        long kh = (np + splits - 1) / splits;
        kh = ((kh + HK - 1) / HK) * HK;
        hipLaunchKernelGGL(contract_h2_kernel, dim3(nwg), dim3(NTH), 0, h->stream, (const float *)pA,
                           (const float *)pB, np, h->Mpad, h->nA2, ntile, splits, kh, sc->eA, sc->eB,
                           sc->p2, sc->n_x, sc->areal, h->d_part);
    } else if (h->contract_3m)
        hipLaunchKernelGGL((contract_kernel<P, true>), dim3(nwg), dim3(NTH), 0, h->stream, (const P *)pA,
                           (const P *)pB, np, h->Mpad, h->nA2, ntile, splits, kper, h->d_part);
    else
        hipLaunchKernelGGL((contract_kernel<P, false>), dim3(nwg), dim3(NTH), 0, h->stream, (const P *)pA,
                           (const P *)pB, np, h->Mpad, h->nA2, ntile, splits, kper, h->d_part);
    TH_HIP(hipGetLastError());
    const size_t n = (size_t)h->nA2 * 2 * h->Mpad * h->Mpad;
    double *acc = h->d_acc + (size_t)axis * n;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                       (const double *)h->d_part, acc, splits, n);
    TH_HIP(hipGetLastError());
    h->n_launch[1] += 1;
    h->n_launch[2] += 1;
    return TH_OK;
}

}  // namespace

int th_pow2_scale(double bound)
{
    if (!(bound > 0.0) || !std::isfinite(bound)) return 0;
    const int E = std::ilogb(bound) + 1;   // bound < 2^E
    return std::min(100, std::max(-100, 14 - E));
}

int th_launch_contract(th_handle *h, const void *pA, const void *pB, long np, int prec, int axis,
                       const th_contract_scale *sc)
{
    if (np <= 0) return TH_OK;
    if (prec == TH_PREC_FP64) return launch_contract_t<double>(h, pA, pB, np, axis, nullptr);
    return launch_contract_t<float>(h, pA, pB, np, axis, sc);
}

int th_launch_zero(th_handle *h, double *p, size_t n)
{
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, p, n);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

int th_launch_finalize(th_handle *h, double2 *chi0)
{
    const size_t n = (size_t)h->D * h->D;
    hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                       (const double *)h->d_acc, chi0, h->nA, h->nOrb, h->Mpad, h->D);
    TH_HIP(hipGetLastError());
    h->n_launch[2] += 1;
    return TH_OK;
}

int th_launch_panel_adjoint(th_handle *h, const void *panel, int prec, const double *That, long np, int trans,
                            double *out, const th_rowmap &rows, int split_e, const float *blkmax, long nblkmax)
{
    const float unsplit = split_e == INT_MIN ? 0.0f : ldexpf(1.0f, -split_e);
    if (np <= 0) return TH_OK;
    const int ntile = h->Mpad / TH_TILE;
    const long nptile = (np + TH_TILE - 1) / TH_TILE;
    const long nwg = nptile * ntile * h->nA2;
    const dim3 grid((unsigned)nwg), blk(NTH);
    if (prec != TH_PREC_FP64 && h->contract_f16) {
        // fp16 matrix cores: That split by th_panel_adjoint_prepare, X scaled by
        // its largest entry (FP32 panels of the batch) or its split exponent
        if (h->pa_that != That) {
            th_set_error("panel adjoint: That was not prepared (th_panel_adjoint_prepare)");
            return TH_E_STATE;
        }
        unsigned *mx = static_cast<unsigned *>(h->d_pamax);
        hipLaunchKernelGGL(zero_u_kernel, dim3(1), dim3(64), 0, h->stream, mx, 1);
        const size_t ps = (size_t)h->nA2 * 2 * h->Mpad;
        if (split_e == INT_MIN && blkmax && nblkmax > 0) {
            // the largest entry from the Green kernel's per-workgroup maxima
            hipLaunchKernelGGL(max_of_kernel, dim3(1), dim3(256), 0, h->stream, blkmax, nblkmax, mx);
        } else if (split_e == INT_MIN) {
            // the rows the batch reads: pg0 .. pg0 + np - 1, or through a row map
            // the whole cache batches that hold them (a wider set, still a bound)
            long r0 = rows.pg0, r1 = rows.pg0 + np;
            if (rows.bpc > 0) {
                r0 = (rows.pg0 / rows.bpc) * rows.bpc;
                r1 = std::min(rows.ntot, ((rows.pg0 + np + rows.bpc - 1) / rows.bpc) * rows.bpc);
            }
            hipLaunchKernelGGL(absmax_f_kernel, dim3(1024), dim3(256), 0, h->stream,
                               (const float *)panel + (size_t)r0 * ps, (size_t)(r1 - r0) * ps, 0.0f, mx);
        }
        const int Kp = ((h->Mpad + PAK - 1) / PAK) * PAK;
        const size_t ny = (size_t)h->nA2 * 2 * h->Mpad * Kp;
        const _Float16 *Ys = static_cast<const _Float16 *>(h->d_pays) + (size_t)trans * ny * 2;
        hipLaunchKernelGGL(panel_adjoint_h2_kernel, grid, dim3(448), 0, h->stream, (const float *)panel,
                           (const _Float16 *)Ys, Kp, np, h->Mpad, h->nA2, ntile, trans, out, rows, unsplit, split_e,
                           (const unsigned *)mx);
        TH_HIP(hipGetLastError());
        return TH_OK;
    }
    if (prec == TH_PREC_FP64) {
        if (h->contract_3m)
            hipLaunchKernelGGL((panel_adjoint_kernel<double, true>), grid, blk, 0, h->stream, (const double *)panel,
                               That, np, h->Mpad, h->nA2, ntile, trans, out, rows, unsplit);
        else
            hipLaunchKernelGGL((panel_adjoint_kernel<double, false>), grid, blk, 0, h->stream, (const double *)panel,
                               That, np, h->Mpad, h->nA2, ntile, trans, out, rows, unsplit);
    } else {
        if (h->contract_3m)
            hipLaunchKernelGGL((panel_adjoint_kernel<float, true>), grid, blk, 0, h->stream, (const float *)panel,
                               That, np, h->Mpad, h->nA2, ntile, trans, out, rows, unsplit);
        else
            hipLaunchKernelGGL((panel_adjoint_kernel<float, false>), grid, blk, 0, h->stream, (const float *)panel,
                               That, np, h->Mpad, h->nA2, ntile, trans, out, rows, unsplit);
    }
    TH_HIP(hipGetLastError());
    return TH_OK;
}

int th_panel_adjoint_prepare(th_handle *h, int prec, const double *That)
{
    if (prec == TH_PREC_FP64 || !h->contract_f16) return TH_OK;
    if (!h->d_pamax) TH_HIP(hipMalloc(&h->d_pamax, 2 * sizeof(unsigned)));
    const int Kp = ((h->Mpad + PAK - 1) / PAK) * PAK;
    const size_t ny = (size_t)h->nA2 * 2 * h->Mpad * Kp;
    if (!h->d_pays) TH_HIP(hipMalloc(&h->d_pays, 2 * ny * 2 * sizeof(_Float16)));
    unsigned *mx = static_cast<unsigned *>(h->d_pamax);
    hipLaunchKernelGGL(zero_u_kernel, dim3(1), dim3(64), 0, h->stream, mx, 2);
    hipLaunchKernelGGL(absmax_d_kernel, dim3(1024), dim3(256), 0, h->stream, That,
                       (size_t)h->nA2 * 2 * h->Mpad * h->Mpad, mx + 1);
    for (int trans = 0; trans < 2; ++trans)
        hipLaunchKernelGGL(presplit_y_kernel, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, h->stream, That,
                           h->Mpad, Kp, h->nA2, trans, (const unsigned *)mx,
                           static_cast<_Float16 *>(h->d_pays) + (size_t)trans * ny * 2);
    TH_HIP(hipGetLastError());
    h->pa_that = That;
    return TH_OK;
}
