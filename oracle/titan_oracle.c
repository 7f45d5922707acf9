/*
 * titan_oracle.c -- FP64 CPU restatement of TITAN's chi(omega) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * titan_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path never calls it.
 *
 * Every function restates the mathematics of one reference routine (cited as
 * file:line into TITAN-FZJ/TITAN source/), written from scratch in C99 with
 * OpenMP.  All matrices are column-major like Fortran: A[i + j*lda].
 * Complex numbers are C99 double complex (interleaved re/im, the same memory
 * layout as Fortran complex(dp)).
 *
 * Pinning (see tests/test_oracle_*.py):
 *   or_gauleg            vs numpy leggauss; polynomial exactness
 *   or_imag/real_epoints vs closed-form integrals of the y=x/(1-x) map
 *   or_intd              Slater-Koster invariants (parity, z-axis limits,
 *                        rotation covariance) + Fe fixture end to end
 *   or_build_system      Hermiticity, periodicity, Fe/nanoribbon fixtures
 *   or_zinv              residual and brute force vs numpy.linalg.inv
 *   or_chi0              Fe_bulk_susc chihf_1_1 fixture (6 omega), nanoribbon
 *                        chihf_i_j fixture, one-orbital Lindhard closed form
 *   or_dyson             Fe_bulk_susc chi_1_1 fixture, scalar closed form
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;

#define OR_MAXORB 9

/* ------------------------------------------------------------------------ */
/* Energy contour                                                            */
/* ------------------------------------------------------------------------ */

/* Gauss-Legendre nodes/weights on [x1,x2] by Newton iteration on P_n
 * starting from the Tricomi guess cos(pi(i-1/4)/(n+1/2)).
 * Follows EnergyIntegration.F90:125-157 (same start guess with the same
 * 10-digit pi and the same 3e-14 stopping rule, so nodes agree to ulps). */
void or_gauleg(double x1, double x2, int n, double *x, double *w)
{
    const double stop = 3.0e-14;
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; ++i) {
        double z = cos(3.141592654 * (i - 0.25) / (n + 0.5));
        double z1 = 0.0, pp = 0.0;
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
            /* recompute pp at the accepted z like the reference does: the
             * reference keeps pp from the last Newton step */
            (void)p2;
        } while (fabs(z - z1) > stop);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = (2.0 * xl) / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* Imaginary-axis nodes y (ascending) and weights for the integral over
 * y in [0,inf) through x in [0,1), y = x/(1-x), dy = dx/(1-x)^2.
 * 'parts' segments: [0.01,1], [1e-(k+1),1e-k] for 1<k<parts, [0,1e-parts].
 * EnergyIntegration.F90:45-97. */
void or_imag_epoints(int parts, int n1gl, double *y, double *wght)
{
    int pn1 = parts * n1gl;
    double *x1 = (double *)malloc(sizeof(double) * pn1);
    double *p1 = (double *)malloc(sizeof(double) * pn1);
    for (int k = parts; k >= 1; --k) {
        double e1, e2;
        if (parts != 1) {
            if (k == 1) { e1 = 1e-2; e2 = 1.0; }
            else if (k != parts) { e1 = pow(10.0, -k - 1); e2 = pow(10.0, -k); }
            else { e1 = 0.0; e2 = pow(10.0, -k); }
        } else { e1 = 0.0; e2 = 1.0; }
        or_gauleg(e1, e2, n1gl, x1 + (k - 1) * n1gl, p1 + (k - 1) * n1gl);
    }
    for (int k = 0; k < pn1; ++k) {
        y[k] = x1[k] / (1.0 - x1[k]);
        wght[k] = p1[k] / ((1.0 - x1[k]) * (1.0 - x1[k]));
    }
    /* ascending order in y (exchange sort as the reference; stable enough
     * for distinct nodes) */
    for (int i = 0; i < pn1; ++i)
        for (int j = i + 1; j < pn1; ++j)
            if (y[i] > y[j]) {
                double t = y[i]; y[i] = y[j]; y[j] = t;
                t = wght[i]; wght[i] = wght[j]; wght[j] = t;
            }
    free(x1); free(p1);
}

/* Real-axis nodes for [Ef-e, Ef] split into parts3 equal segments.
 * EnergyIntegration.F90:99-122 (only used when |e| >= 1e-15). */
void or_real_epoints(double Ef, double e, int parts3, int n3gl, double *x2, double *p2)
{
    for (int k = 1; k <= parts3; ++k) {
        double e1 = Ef - ((parts3 - (k - 1)) * e / parts3);
        double e2 = Ef - ((parts3 - k) * e / parts3);
        or_gauleg(e1, e2, n3gl, x2 + (k - 1) * n3gl, p2 + (k - 1) * n3gl);
    }
}

/* ------------------------------------------------------------------------ */
/* Slater-Koster two-centre table                                            */
/* ------------------------------------------------------------------------ */

/* 9x9 two-centre hopping table b(row,col) (column-major, 0-based orbital
 * ids s,px,py,pz,dxy,dyz,dzx,dx2-y2,dz2) for direction cosines w=(l,m,n)
 * and the ten integrals p = (sss,pps,ppp,dds,ddp,ddd,sps,sds,pds,pdp).
 * Slater & Koster, Phys. Rev. 94, 1498 (1954), Table I; orbital order and
 * sign conventions as TightBinding.F90:655-783 (b(j,i) = (-1)^(l+l') b(i,j)).
 * Written here from the published table, one matrix element per line. */
void or_intd(const double *p, const double *w, double *b)
{
    const double sss = p[0], pps = p[1], ppp = p[2], dds = p[3], ddp = p[4],
                 ddd = p[5], sps = p[6], sds = p[7], pds = p[8], pdp = p[9];
    const double l = w[0], m = w[1], n = w[2];
    const double l2 = l * l, m2 = m * m, n2 = n * n;
    const double r3 = sqrt(3.0);
    double E[9][9];
    /* s row */
    E[0][0] = sss;
    E[0][1] = l * sps;
    E[0][2] = m * sps;
    E[0][3] = n * sps;
    E[0][4] = r3 * l * m * sds;
    E[0][5] = r3 * m * n * sds;
    E[0][6] = r3 * n * l * sds;
    E[0][7] = 0.5 * r3 * (l2 - m2) * sds;
    E[0][8] = (n2 - 0.5 * (l2 + m2)) * sds;
    /* p-p block */
    E[1][1] = l2 * pps + (1.0 - l2) * ppp;
    E[2][2] = m2 * pps + (1.0 - m2) * ppp;
    E[3][3] = n2 * pps + (1.0 - n2) * ppp;
    E[1][2] = l * m * (pps - ppp);
    E[1][3] = l * n * (pps - ppp);
    E[2][3] = m * n * (pps - ppp);
    /* p-d block: x row */
    E[1][4] = r3 * l2 * m * pds + m * (1.0 - 2.0 * l2) * pdp;
    E[1][5] = r3 * l * m * n * pds - 2.0 * l * m * n * pdp;
    E[1][6] = r3 * l2 * n * pds + n * (1.0 - 2.0 * l2) * pdp;
    E[1][7] = 0.5 * r3 * l * (l2 - m2) * pds + l * (1.0 - l2 + m2) * pdp;
    E[1][8] = l * (n2 - 0.5 * (l2 + m2)) * pds - r3 * l * n2 * pdp;
    /* y row */
    E[2][4] = r3 * m2 * l * pds + l * (1.0 - 2.0 * m2) * pdp;
    E[2][5] = r3 * m2 * n * pds + n * (1.0 - 2.0 * m2) * pdp;
    E[2][6] = r3 * l * m * n * pds - 2.0 * l * m * n * pdp;
    E[2][7] = 0.5 * r3 * m * (l2 - m2) * pds - m * (1.0 + l2 - m2) * pdp;
    E[2][8] = m * (n2 - 0.5 * (l2 + m2)) * pds - r3 * m * n2 * pdp;
    /* z row */
    E[3][4] = r3 * l * m * n * pds - 2.0 * l * m * n * pdp;
    E[3][5] = r3 * n2 * m * pds + m * (1.0 - 2.0 * n2) * pdp;
    E[3][6] = r3 * n2 * l * pds + l * (1.0 - 2.0 * n2) * pdp;
    E[3][7] = 0.5 * r3 * n * (l2 - m2) * pds - n * (l2 - m2) * pdp;
    E[3][8] = n * (n2 - 0.5 * (l2 + m2)) * pds + r3 * n * (l2 + m2) * pdp;
    /* d-d block */
    E[4][4] = 3.0 * l2 * m2 * dds + (l2 + m2 - 4.0 * l2 * m2) * ddp + (n2 + l2 * m2) * ddd;
    E[5][5] = 3.0 * m2 * n2 * dds + (m2 + n2 - 4.0 * m2 * n2) * ddp + (l2 + m2 * n2) * ddd;
    E[6][6] = 3.0 * n2 * l2 * dds + (n2 + l2 - 4.0 * n2 * l2) * ddp + (m2 + n2 * l2) * ddd;
    E[4][5] = 3.0 * l * m2 * n * dds + l * n * (1.0 - 4.0 * m2) * ddp + l * n * (m2 - 1.0) * ddd;
    E[4][6] = 3.0 * l2 * m * n * dds + m * n * (1.0 - 4.0 * l2) * ddp + m * n * (l2 - 1.0) * ddd;
    E[5][6] = 3.0 * m * n2 * l * dds + m * l * (1.0 - 4.0 * n2) * ddp + m * l * (n2 - 1.0) * ddd;
    E[4][7] = 1.5 * l * m * (l2 - m2) * dds + 2.0 * l * m * (m2 - l2) * ddp + 0.5 * l * m * (l2 - m2) * ddd;
    E[5][7] = 1.5 * m * n * (l2 - m2) * dds - m * n * (1.0 + 2.0 * (l2 - m2)) * ddp + m * n * (1.0 + 0.5 * (l2 - m2)) * ddd;
    E[6][7] = 1.5 * n * l * (l2 - m2) * dds + n * l * (1.0 - 2.0 * (l2 - m2)) * ddp - n * l * (1.0 - 0.5 * (l2 - m2)) * ddd;
    E[4][8] = r3 * l * m * (n2 - 0.5 * (l2 + m2)) * dds - 2.0 * r3 * l * m * n2 * ddp + 0.5 * r3 * l * m * (1.0 + n2) * ddd;
    E[5][8] = r3 * m * n * (n2 - 0.5 * (l2 + m2)) * dds + r3 * m * n * (l2 + m2 - n2) * ddp - 0.5 * r3 * m * n * (l2 + m2) * ddd;
    E[6][8] = r3 * l * n * (n2 - 0.5 * (l2 + m2)) * dds + r3 * l * n * (l2 + m2 - n2) * ddp - 0.5 * r3 * l * n * (l2 + m2) * ddd;
    E[7][7] = 0.75 * (l2 - m2) * (l2 - m2) * dds + (l2 + m2 - (l2 - m2) * (l2 - m2)) * ddp
              + (n2 + 0.25 * (l2 - m2) * (l2 - m2)) * ddd;
    E[7][8] = 0.5 * r3 * (l2 - m2) * (n2 - 0.5 * (l2 + m2)) * dds + r3 * n2 * (m2 - l2) * ddp
              + 0.25 * r3 * (1.0 + n2) * (l2 - m2) * ddd;
    E[8][8] = (n2 - 0.5 * (l2 + m2)) * (n2 - 0.5 * (l2 + m2)) * dds + 3.0 * n2 * (l2 + m2) * ddp
              + 0.75 * (l2 + m2) * (l2 + m2) * ddd;
    /* lower triangle by the parity rule E(b,a) = (-1)^(l_a+l_b) E(a,b) */
    static const int ang[9] = {0, 1, 1, 1, 2, 2, 2, 2, 2};
    for (int a = 0; a < 9; ++a)
        for (int c = a + 1; c < 9; ++c)
            E[c][a] = ((ang[a] + ang[c]) & 1) ? -E[a][c] : E[a][c];
    for (int a = 0; a < 9; ++a)
        for (int c = 0; c < 9; ++c)
            b[a + 9 * c] = E[a][c];
}

/* ------------------------------------------------------------------------ */
/* Orbital angular momentum in the cubic-harmonic basis                      */
/* ------------------------------------------------------------------------ */

/* L_x, L_y, L_z as 9x9 complex matrices (column-major) in the orbital order
 * of or_intd.  Matrix elements as mod_magnet.F90:295-358 (L+ table and
 * Lz, then Lx = (L+ + L-)/2, Ly = -i(L+ - L-)/2 with L- = (L+)^dagger). */
void or_lmat(zc *L /* 3*81 */)
{
    zc Lp[81], Lz[81];
    const double r3 = sqrt(3.0);
    memset(Lp, 0, sizeof(Lp));
    memset(Lz, 0, sizeof(Lz));
#define E9(M, r, c) M[(r - 1) + 9 * (c - 1)]
    E9(Lz, 2, 3) = -I;       E9(Lz, 3, 2) = I;
    E9(Lz, 5, 8) = 2.0 * I;  E9(Lz, 8, 5) = -2.0 * I;
    E9(Lz, 6, 7) = I;        E9(Lz, 7, 6) = -I;
    E9(Lp, 2, 4) = -1.0;     E9(Lp, 3, 4) = -I;
    E9(Lp, 4, 2) = 1.0;      E9(Lp, 4, 3) = I;
    E9(Lp, 5, 6) = -1.0;     E9(Lp, 5, 7) = -I;
    E9(Lp, 6, 5) = 1.0;      E9(Lp, 6, 8) = -I;      E9(Lp, 6, 9) = -r3 * I;
    E9(Lp, 7, 5) = I;        E9(Lp, 7, 8) = 1.0;     E9(Lp, 7, 9) = -r3;
    E9(Lp, 8, 6) = I;        E9(Lp, 8, 7) = -1.0;
    E9(Lp, 9, 6) = r3 * I;   E9(Lp, 9, 7) = r3;
#undef E9
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) {
            zc lp = Lp[r + 9 * c], lm = conj(Lp[c + 9 * r]);
            L[r + 9 * c] = 0.5 * (lp + lm);
            L[81 + r + 9 * c] = -0.5 * I * (lp - lm);
            L[162 + r + 9 * c] = Lz[r + 9 * c];
        }
}

/* ------------------------------------------------------------------------ */
/* System description and set-up                                             */
/* ------------------------------------------------------------------------ */

/* Raw inputs of one tight-binding cell (the numbers TITAN reads from its
 * input, element, basis and self-consistency files).  Energies in Ry,
 * lengths in any unit (only ratios and dot products with k enter). */
typedef struct {
    int nA;              /* atoms in the cell */
    int nOrb;            /* orbitals per atom (same selection for all) */
    int orbs[9];         /* selected orbital ids, 1..9 = s,px,py,pz,dxy,dyz,dzx,dx2,dz2 */
    int sysdim;          /* 1, 2 or 3 */
    int nStages;         /* neighbour shells */
    double relTol;       /* shell tolerance relative to the first shell */
    double a[9];         /* a1,a2,a3 (a[0..2] = a1) */
    const double *pos;   /* 3*nA basis positions */
    const int *type;     /* nA type index, 0-based */
    int nTypes;
    const double *onsite;  /* nTypes*9 on-site energies by orbital id */
    const double *sk;      /* nTypes*nStages*10 two-centre integrals */
    const double *lambda;  /* nTypes*2 : lambda_p, lambda_d */
    const double *U;       /* nTypes*2 : Un, Um */
    const double *rho, *mx, *my, *mz;  /* nA*nOrb ground-state occupations/moments */
    const double *rho0;    /* nA*nOrb reference occupations (0 when Un=0) */
    const double *rhod0, *mzd0, *mpd0_re, *mpd0_im; /* nA reference d sums */
    double socscale;
    int simplemix;         /* 1: arithmetic mean of the two types' integrals */
    /* Andersen scaling (get_SK_parameter, TightBinding.F90:144-162): nel[t]
     * element-file atoms of type t's own element, each with its nStages shell
     * distances estage[((t*maxel) + l)*nStages + k]; nel NULL or nel[t] = 0:
     * no scaling of type t */
    int maxel;
    const int *nel;
    const double *estage;
} or_sys_in;

/* Hop table: H(k)[row block jb, col block ib] += exp(i k.R) t (both spins).
 * t is nOrb x nOrb real (column-major, rows = orbitals of jb). */
typedef struct {
    int nA, nOrb, d, nhop;
    double *R;     /* 3*nhop cell vectors */
    int *jb, *ib;  /* nhop */
    double *t;     /* nhop*nOrb*nOrb */
    zc *h0;        /* d*d local Hamiltonian */
    int D;         /* 4*nOrb*nOrb*nA */
    zc *Umat;      /* D*D interaction matrix in the chi basis */
    double b[9];   /* reciprocal basis b1,b2,b3 (b[0..2] = b1) */
} or_sys;

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Row index of (atom i, spin s in {0=up,1=down}, orbital mu) in H:
 * atom-major, then spin, then orbital (mod_System.F90:101-155, 'ia'). */
static inline int hidx(int nOrb, int i, int s, int mu) { return (2 * i + s) * nOrb + mu; }

/* chi-basis index (sigma in 0..3, atom i, mu, nu), sigma fastest, then nu,
 * then mu, then atom (mod_System.F90:158-223, 'sigmaimunu2i'). */
static inline int cidx(int nOrb, int sigma, int i, int mu, int nu)
{
    return 4 * ((i * nOrb + mu) * nOrb + nu) + sigma;
}

static double vdist(const double *a, const double *b)
{
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* Builds the hop table, h0 and the interaction matrix from raw inputs.
 * Neighbour shells: Lattice.F90:40-232; hoppings: TightBinding.F90:93-219
 * (geometric-mean mixing of the two types' integrals, no distance scaling:
 * the element lattice equals the system lattice for every deck used here);
 * h0: mod_hamiltonian.F90:35-87 with SOC from mod_SOC.F90:13-74 (theta=phi=0)
 * and the mean-field term of mod_U_matrix.F90:29-55; Umat:
 * mod_susceptibilities.F90:106-159.  Returns 0 on success. */
int or_build_system(const or_sys_in *in, or_sys *out)
{
    const int nA = in->nA, nO = in->nOrb, nS = in->nStages;
    const int d = 2 * nO * nA;
    int ncell1 = 2 * nS + 1;
    int nCells = 1;
    for (int k = 0; k < in->sysdim; ++k) nCells *= ncell1;
    int isize = nCells * nA;
    double *cellv = (double *)malloc(sizeof(double) * 3 * isize);
    double *posl = (double *)malloc(sizeof(double) * 3 * isize);
    int *bidx = (int *)malloc(sizeof(int) * isize);
    double *dist = (double *)malloc(sizeof(double) * isize * nA);
    double *sorted = (double *)malloc(sizeof(double) * isize * nA);
    double *stages = (double *)malloc(sizeof(double) * nS * nA);
    int cnt = 0;
    for (int c = 0; c < nCells; ++c) {
        int n1 = c % ncell1 - nS, n2 = 0, n3 = 0;
        if (in->sysdim >= 2) n2 = (c / ncell1) % ncell1 - nS;
        if (in->sysdim >= 3) n3 = (c / (ncell1 * ncell1)) % ncell1 - nS;
        for (int j = 0; j < nA; ++j, ++cnt) {
            for (int x = 0; x < 3; ++x) {
                cellv[3 * cnt + x] = n1 * in->a[x] + n2 * in->a[3 + x] + n3 * in->a[6 + x];
                posl[3 * cnt + x] = in->pos[3 * j + x] + cellv[3 * cnt + x];
            }
            bidx[cnt] = j;
            for (int k = 0; k < nA; ++k) dist[cnt * nA + k] = vdist(posl + 3 * cnt, in->pos + 3 * k);
        }
    }
    /* distinct shell radii per cell-0 atom */
    for (int k = 0; k < nA; ++k) {
        int m = 0;
        for (int i = 0; i < isize; ++i)
            if (dist[i * nA + k] > 1e-9) sorted[m++] = dist[i * nA + k];
        /* insertion sort ascending */
        for (int i = 1; i < m; ++i) {
            double v = sorted[i];
            int l = i - 1;
            while (l >= 0 && sorted[l] > v) { sorted[l + 1] = sorted[l]; --l; }
            sorted[l + 1] = v;
        }
        for (int s = 0; s < nS; ++s) stages[s * nA + k] = 1e12;
        stages[0 * nA + k] = sorted[0];
        int l = 0;
        for (int i = 1; i < m; ++i) {
            if (fabs(sorted[i] - stages[l * nA + k]) < stages[0 * nA + k] * in->relTol) continue;
            ++l;
            if (l >= nS) break;
            stages[l * nA + k] = sorted[i];
        }
    }
    /* hop table: one entry per (neighbour site, cell-0 atom it hops to) */
    int cap = isize * nA;
    out->R = (double *)malloc(sizeof(double) * 3 * cap);
    out->jb = (int *)malloc(sizeof(int) * cap);
    out->ib = (int *)malloc(sizeof(int) * cap);
    out->t = (double *)malloc(sizeof(double) * cap * nO * nO);
    int nh = 0;
    for (int s = 0; s < isize; ++s) {
        int j = bidx[s];
        for (int i = 0; i < nA; ++i) {
            double dd = dist[s * nA + i];
            if (dd <= 1e-9) continue;
            int stage = -1;
            for (int k = 0; k < nS; ++k)
                if (fabs(dd - stages[k * nA + i]) < stages[0 * nA + i] * in->relTol) { stage = k; break; }
            if (stage < 0) continue;
            double dc[3];
            for (int x = 0; x < 3; ++x) dc[x] = (posl[3 * s + x] - in->pos[3 * i + x]) / dd;
            const double *t1 = in->sk + ((size_t)in->type[i] * nS + stage) * 10;
            const double *t2 = in->sk + ((size_t)in->type[j] * nS + stage) * 10;
            /* distance exponents l + l' + 1 of the ten integrals */
            static const double expon[10] = {1, 3, 3, 5, 5, 5, 2, 3, 4, 4};
            double sc[2] = {1.0, 1.0};
            for (int side = 0; side < 2; ++side) {
                const int ty = in->type[side ? j : i];
                if (!in->nel || in->nel[ty] <= 0) continue;
                double best = 0.0;   /* element atom whose shell ratio is closest to 1 */
                for (int l = 0; l < in->nel[ty]; ++l) {
                    double r = in->estage[((size_t)ty * in->maxel + l) * nS + stage] / dd;
                    if (fabs(r - 1.0) < fabs(best - 1.0)) best = r;
                }
                sc[side] = best;
            }
            double mix[10], bb[81];
            for (int p = 0; p < 10; ++p) {
                const double a1 = t1[p] * pow(sc[0], expon[p]), a2 = t2[p] * pow(sc[1], expon[p]);
                if (in->simplemix) {
                    mix[p] = 0.5 * (a1 + a2);
                } else {
                    double g = sqrt(fabs(a1) * fabs(a2));
                    mix[p] = (a1 + a2 >= 0.0) ? g : -g;  /* Fortran sign(a,b) */
                }
            }
            or_intd(mix, dc, bb);
            for (int x = 0; x < 3; ++x) out->R[3 * nh + x] = cellv[3 * s + x];
            out->jb[nh] = j;
            out->ib[nh] = i;
            for (int nu = 0; nu < nO; ++nu)
                for (int mu = 0; mu < nO; ++mu)
                    out->t[(size_t)nh * nO * nO + nu + nO * mu] = bb[(in->orbs[nu] - 1) + 9 * (in->orbs[mu] - 1)];
            ++nh;
        }
    }
    out->nhop = nh;
    out->nA = nA; out->nOrb = nO; out->d = d;
    /* h0 */
    out->h0 = (zc *)calloc((size_t)d * d, sizeof(zc));
    zc L[243];
    or_lmat(L);
    for (int i = 0; i < nA; ++i) {
        int ty = in->type[i];
        const double lp = in->socscale * in->lambda[2 * ty], ld = in->socscale * in->lambda[2 * ty + 1];
        const double Un = in->U[2 * ty], Um = in->U[2 * ty + 1];
        for (int mu = 0; mu < nO; ++mu) {
            double e = in->onsite[ty * 9 + in->orbs[mu] - 1];
            out->h0[hidx(nO, i, 0, mu) + d * hidx(nO, i, 0, mu)] += e;
            out->h0[hidx(nO, i, 1, mu) + d * hidx(nO, i, 1, mu)] += e;
        }
        /* SOC: (1/2) lambda L.sigma within the p block and the d block */
        for (int mu = 0; mu < nO; ++mu)
            for (int nu = 0; nu < nO; ++nu) {
                int om = in->orbs[mu], on = in->orbs[nu];
                double lam;
                if (om >= 2 && om <= 4 && on >= 2 && on <= 4) lam = lp;
                else if (om >= 5 && on >= 5) lam = ld;
                else continue;
                zc lx = L[(om - 1) + 9 * (on - 1)], ly = L[81 + (om - 1) + 9 * (on - 1)],
                   lz = L[162 + (om - 1) + 9 * (on - 1)];
                /* sigma_x, sigma_y, sigma_z entries (s,s') */
                zc uu = lz, dn = -lz, ud = lx - I * ly, du = lx + I * ly;
                out->h0[hidx(nO, i, 0, mu) + d * hidx(nO, i, 0, nu)] += 0.5 * lam * uu;
                out->h0[hidx(nO, i, 1, mu) + d * hidx(nO, i, 1, nu)] += 0.5 * lam * dn;
                out->h0[hidx(nO, i, 0, mu) + d * hidx(nO, i, 1, nu)] += 0.5 * lam * ud;
                out->h0[hidx(nO, i, 1, mu) + d * hidx(nO, i, 0, nu)] += 0.5 * lam * du;
            }
        /* mean-field exchange on d orbitals */
        double rhod = 0, mxd = 0, myd = 0, mzd = 0;
        for (int mu = 0; mu < nO; ++mu)
            if (in->orbs[mu] >= 5) {
                rhod += in->rho[i * nO + mu];
                mxd += in->mx[i * nO + mu];
                myd += in->my[i * nO + mu];
                mzd += in->mz[i * nO + mu];
            }
        zc mpd = mxd + I * myd, mpd0 = in->mpd0_re[i] + I * in->mpd0_im[i];
        for (int mu = 0; mu < nO; ++mu) {
            if (in->orbs[mu] < 5) continue;
            double dn = Un * (in->rho[i * nO + mu] - in->rho0[i * nO + mu]) + 0.5 * Un * (-rhod + in->rhod0[i]);
            out->h0[hidx(nO, i, 0, mu) + d * hidx(nO, i, 0, mu)] += dn + 0.5 * Um * (-mzd + in->mzd0[i]);
            out->h0[hidx(nO, i, 1, mu) + d * hidx(nO, i, 1, mu)] += dn + 0.5 * Um * (mzd - in->mzd0[i]);
            out->h0[hidx(nO, i, 0, mu) + d * hidx(nO, i, 1, mu)] += -0.5 * Um * conj(mpd - mpd0);
            out->h0[hidx(nO, i, 1, mu) + d * hidx(nO, i, 0, mu)] += -0.5 * Um * (mpd - mpd0);
        }
    }
    /* interaction matrix in the chi basis */
    int D = 4 * nO * nO * nA;
    out->D = D;
    out->Umat = (zc *)calloc((size_t)D * D, sizeof(zc));
    for (int i = 0; i < nA; ++i) {
        int ty = in->type[i];
        double Un = in->U[2 * ty], Um = in->U[2 * ty + 1];
        for (int mu = 0; mu < nO; ++mu) {
            if (in->orbs[mu] < 5) continue;
            for (int ga = 0; ga < nO; ++ga) {
                if (in->orbs[ga] < 5) continue;
                out->Umat[cidx(nO, 0, i, mu, mu) + (size_t)D * cidx(nO, 0, i, ga, ga)] = Um;
                out->Umat[cidx(nO, 1, i, mu, mu) + (size_t)D * cidx(nO, 1, i, ga, ga)] = Un;
                out->Umat[cidx(nO, 2, i, mu, mu) + (size_t)D * cidx(nO, 2, i, ga, ga)] = Un;
                out->Umat[cidx(nO, 3, i, mu, mu) + (size_t)D * cidx(nO, 3, i, ga, ga)] = Um;
            }
        }
        for (int mu = 0; mu < nO; ++mu) {
            if (in->orbs[mu] < 5) continue;
            for (int nu = 0; nu < nO; ++nu) {
                if (in->orbs[nu] < 5) continue;
                /* entries (mu,nu),(gama=nu, xi=mu) */
                for (int s1 = 1; s1 <= 2; ++s1)
                    for (int s2 = 1; s2 <= 2; ++s2)
                        out->Umat[cidx(nO, s1, i, mu, nu) + (size_t)D * cidx(nO, s2, i, nu, mu)] -= Un;
            }
        }
    }
    /* reciprocal basis */
    double b[9] = {0};
    const double *a1 = in->a, *a2 = in->a + 3, *a3 = in->a + 6;
    const double tpi = 2.0 * acos(-1.0);
    if (in->sysdim == 3) {
        double c[3];
        cross3(a2, a3, c);
        double vol = tpi / (a1[0] * c[0] + a1[1] * c[1] + a1[2] * c[2]);
        cross3(a2, a3, b); cross3(a3, a1, b + 3); cross3(a1, a2, b + 6);
        for (int x = 0; x < 9; ++x) b[x] *= vol;
    } else if (in->sysdim == 2) {
        /* in-plane duals: b_i . a_j = 2 pi delta_ij */
        double z[3], c1[3], c2[3];
        cross3(a1, a2, z);
        double nz = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
        for (int x = 0; x < 3; ++x) z[x] /= nz;
        cross3(a2, z, c1);
        cross3(z, a1, c2);
        double v1 = a1[0] * c1[0] + a1[1] * c1[1] + a1[2] * c1[2];
        double v2 = a2[0] * c2[0] + a2[1] * c2[1] + a2[2] * c2[2];
        for (int x = 0; x < 3; ++x) { b[x] = tpi * c1[x] / v1; b[3 + x] = tpi * c2[x] / v2; }
    } else {
        double n2 = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
        for (int x = 0; x < 3; ++x) b[x] = tpi * a1[x] / n2;
    }
    memcpy(out->b, b, sizeof(b));
    free(cellv); free(posl); free(bidx); free(dist); free(sorted); free(stages);
    return 0;
}

void or_free_system(or_sys *s)
{
    free(s->R); free(s->jb); free(s->ib); free(s->t); free(s->h0); free(s->Umat);
    memset(s, 0, sizeof(*s));
}

/* A system from explicit tables (hops t real, nhop*nOrb*nOrb column-major,
 * h0 d*d), for derivative checks that perturb the tables themselves.  The
 * interaction matrix is left zero (or_dyson takes U separately). */
int or_sys_from_arrays(int nA, int nOrb, int nhop, const double *R, const int *jb, const int *ib,
                       const double *t, const zc *h0, const double *b, or_sys *out)
{
    memset(out, 0, sizeof(*out));
    out->nA = nA; out->nOrb = nOrb; out->d = 2 * nOrb * nA; out->nhop = nhop;
    out->D = 4 * nOrb * nOrb * nA;
    const int d = out->d;
    out->R = malloc(sizeof(double) * 3 * (nhop > 0 ? nhop : 1));
    out->jb = malloc(sizeof(int) * (nhop > 0 ? nhop : 1));
    out->ib = malloc(sizeof(int) * (nhop > 0 ? nhop : 1));
    out->t = malloc(sizeof(double) * (nhop > 0 ? nhop : 1) * nOrb * nOrb);
    out->h0 = malloc(sizeof(zc) * d * d);
    out->Umat = calloc((size_t)out->D * out->D, sizeof(zc));
    if (!out->R || !out->jb || !out->ib || !out->t || !out->h0 || !out->Umat) return -1;
    memcpy(out->R, R, sizeof(double) * 3 * nhop);
    memcpy(out->jb, jb, sizeof(int) * nhop);
    memcpy(out->ib, ib, sizeof(int) * nhop);
    memcpy(out->t, t, sizeof(double) * nhop * nOrb * nOrb);
    memcpy(out->h0, h0, sizeof(zc) * d * d);
    memcpy(out->b, b, sizeof(double) * 9);
    return 0;
}

/* Copies of the built tables for inspection from Python. */
int or_sys_nhop(const or_sys *s) { return s->nhop; }
void or_sys_export(const or_sys *s, double *R, int *jb, int *ib, double *t, zc *h0, zc *Umat, double *b)
{
    int nO = s->nOrb;
    if (R) memcpy(R, s->R, sizeof(double) * 3 * s->nhop);
    if (jb) memcpy(jb, s->jb, sizeof(int) * s->nhop);
    if (ib) memcpy(ib, s->ib, sizeof(int) * s->nhop);
    if (t) memcpy(t, s->t, sizeof(double) * s->nhop * nO * nO);
    if (h0) memcpy(h0, s->h0, sizeof(zc) * s->d * s->d);
    if (Umat) memcpy(Umat, s->Umat, sizeof(zc) * (size_t)s->D * s->D);
    if (b) memcpy(b, s->b, sizeof(double) * 9);
}

/* ------------------------------------------------------------------------ */
/* H(k), inverse and G                                                       */
/* ------------------------------------------------------------------------ */

/* H(k) = h0 + sum_n exp(i k.R_n) t_n (x) 1_spin.  mod_hamiltonian.F90:183-243 */
void or_hk(const or_sys *s, const double *k, zc *H)
{
    const int d = s->d, nO = s->nOrb;
    memcpy(H, s->h0, sizeof(zc) * d * d);
    for (int n = 0; n < s->nhop; ++n) {
        const double *R = s->R + 3 * n;
        double ph = k[0] * R[0] + k[1] * R[1] + k[2] * R[2];
        zc e = cos(ph) + I * sin(ph);
        int j = s->jb[n], i = s->ib[n];
        const double *t = s->t + (size_t)n * nO * nO;
        for (int sp = 0; sp < 2; ++sp)
            for (int mu = 0; mu < nO; ++mu)
                for (int nu = 0; nu < nO; ++nu)
                    H[hidx(nO, j, sp, nu) + d * hidx(nO, i, sp, mu)] += t[nu + nO * mu] * e;
    }
}

static inline double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* In-place inverse by LU with partial pivoting (pivot = largest |re|+|im|),
 * then inverse of U and the triangular solve, as LAPACK zgetrf+zgetri
 * (mod_tools.F90:796-829).  Returns 0, or k>0 if U(k,k) is exactly zero. */
int or_zinv(int n, zc *A)
{
    int *piv = (int *)malloc(sizeof(int) * n);
    zc *work = (zc *)malloc(sizeof(zc) * n);
    int info = 0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        double best = cabs1(A[c + n * c]);
        for (int r = c + 1; r < n; ++r) {
            double v = cabs1(A[r + n * c]);
            if (v > best) { best = v; p = r; }
        }
        piv[c] = p;
        if (best == 0.0) { if (!info) info = c + 1; continue; }
        if (p != c)
            for (int j = 0; j < n; ++j) { zc t = A[c + n * j]; A[c + n * j] = A[p + n * j]; A[p + n * j] = t; }
        zc inv = 1.0 / A[c + n * c];
        for (int r = c + 1; r < n; ++r) A[r + n * c] *= inv;
        for (int j = c + 1; j < n; ++j) {
            zc f = A[c + n * j];
            if (f != 0.0)
                for (int r = c + 1; r < n; ++r) A[r + n * j] -= A[r + n * c] * f;
        }
    }
    if (info) { free(piv); free(work); return info; }
    /* inv(U) in place (upper triangle) */
    for (int j = 0; j < n; ++j) {
        A[j + n * j] = 1.0 / A[j + n * j];
        zc ajj = -A[j + n * j];
        for (int r = 0; r < j; ++r) {
            zc acc = 0.0;
            for (int k = r; k < j; ++k) acc += A[r + n * k] * A[k + n * j];
            work[r] = acc;
        }
        for (int r = 0; r < j; ++r) A[r + n * j] = work[r] * ajj;
    }
    /* solve inv(A) L = inv(U) for inv(A), columns right to left */
    for (int j = n - 1; j >= 0; --j) {
        for (int r = j + 1; r < n; ++r) { work[r] = A[r + n * j]; A[r + n * j] = 0.0; }
        for (int r = j + 1; r < n; ++r) {
            zc f = work[r];
            if (f != 0.0)
                for (int i = 0; i < n; ++i) A[i + n * j] -= A[i + n * r] * f;
        }
    }
    for (int j = n - 2; j >= 0; --j) {
        int p = piv[j];
        if (p != j)
            for (int i = 0; i < n; ++i) { zc t = A[i + n * j]; A[i + n * j] = A[i + n * p]; A[i + n * p] = t; }
    }
    free(piv); free(work);
    return 0;
}

/* G(k,z) = (z - H(k))^-1, z = er + i ei.  mod_greenfunction.F90:19-75. */
int or_green(const or_sys *s, double er, double ei, const double *k, zc *G)
{
    const int d = s->d;
    or_hk(s, k, G);
    for (int i = 0; i < d * d; ++i) G[i] = -G[i];
    for (int i = 0; i < d; ++i) G[i + d * i] += er + I * ei;
    return or_zinv(d, G);
}

/* ------------------------------------------------------------------------ */
/* k meshes                                                                  */
/* ------------------------------------------------------------------------ */

/* Plain uniform mesh k = sum_i n_i b_i / N_i, weight 1/prod(N).  TITAN
 * folds the same points into the first zone and splits boundary duplicates
 * (mod_BrillouinZone.F90:151-238); for H(k) periodic in the reciprocal
 * lattice both meshes give identical sums (checked by or_kmesh_titan). */
void or_kmesh_uniform(const or_sys *s, int n1, int n2, int n3, double *kp, double *w)
{
    int nk = n1 * n2 * n3, c = 0;
    for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n2; ++b)
            for (int e = 0; e < n3; ++e, ++c) {
                for (int x = 0; x < 3; ++x)
                    kp[3 * c + x] = (double)a * s->b[x] / n1 + (double)b * s->b[3 + x] / n2 + (double)e * s->b[6 + x] / n3;
                w[c] = 1.0 / nk;
            }
}

/* TITAN's 3D mesh with first-zone folding and fractional weights for
 * boundary duplicates (mod_BrillouinZone.F90:151-313).  Returns the number
 * of points written (call with kp=NULL to count). */
long or_kmesh_titan3d(const or_sys *s, int N, double *kp, double *w)
{
    const double *b1 = s->b, *b2 = s->b + 3, *b3 = s->b + 6;
    double bz[8][3];
    for (int x = 0; x < 3; ++x) {
        bz[0][x] = 0; bz[1][x] = b1[x]; bz[2][x] = b2[x]; bz[3][x] = b1[x] + b2[x];
        bz[4][x] = b3[x]; bz[5][x] = b1[x] + b3[x]; bz[6][x] = b2[x] + b3[x]; bz[7][x] = b1[x] + b2[x] + b3[x];
    }
    long nk = (long)N * N * N, cnt = 0;
    for (long l = 0; l < nk; ++l) {
        int nx = (int)((l / ((long)N * N)) % N), ny = (int)((l / N) % N), nz = (int)(l % N);
        double k[3], diff[8][3], dist[8], best = 1e300;
        for (int x = 0; x < 3; ++x) k[x] = nx * b1[x] / N + ny * b2[x] / N + nz * b3[x] / N;
        for (int j = 0; j < 8; ++j) {
            for (int x = 0; x < 3; ++x) diff[j][x] = k[x] - bz[j][x];
            dist[j] = sqrt(diff[j][0] * diff[j][0] + diff[j][1] * diff[j][1] + diff[j][2] * diff[j][2]);
            if (dist[j] < best) best = dist[j];
        }
        int mult = 0;
        for (int j = 0; j < 8; ++j) if (fabs(dist[j] - best) < 1e-12) ++mult;
        for (int j = 0; j < 8; ++j)
            if (fabs(dist[j] - best) < 1e-12) {
                if (kp) {
                    for (int x = 0; x < 3; ++x) kp[3 * cnt + x] = diff[j][x];
                    w[cnt] = 1.0 / mult / (double)nk;
                }
                ++cnt;
            }
    }
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* chi0: the eintshechi contour integral                                     */
/* ------------------------------------------------------------------------ */

/* Spin block extraction: G(atom i, spin a, orb mu ; atom j, spin b, orb nu) */
#define GB(G, d, nO, i, a, mu, j, b, nu) ((G)[hidx(nO, i, a, mu) + (d) * hidx(nO, j, b, nu)])

/* Contribution of one imaginary-axis (k,E) pair with weight w into acc (D x D):
 * the 16 spin combinations of eintshechi.F90:83-105, with Gq = G(k+q,
 * E_F+omega+i(y+eta)) and G = G(k, E_F+i(y+eta')).  Spin labels u=0, d=1. */
static void chi0_imag_pair(int nA, int nO, int d, int D, const zc *Gq, const zc *G, double w, zc *acc)
{
    /* (a_sigma, c_sigma) for sigma = +,up,down,- : a = (d,u,d,u), c = (u,u,d,d) */
    static const int A_[4] = {1, 0, 1, 0}, C_[4] = {0, 0, 1, 1};
    for (int j = 0; j < nA; ++j)
        for (int i = 0; i < nA; ++i)
            for (int ga = 0; ga < nO; ++ga)
                for (int mu = 0; mu < nO; ++mu)
                    for (int xi = 0; xi < nO; ++xi)
                        for (int nu = 0; nu < nO; ++nu) {
                            for (int s1 = 0; s1 < 4; ++s1)
                                for (int s2 = 0; s2 < 4; ++s2) {
                                    zc t1 = GB(Gq, d, nO, i, A_[s1], nu, j, A_[s2], ga) *
                                            GB(G, d, nO, j, C_[s2], xi, i, C_[s1], mu);
                                    zc t2 = GB(G, d, nO, i, C_[s1], mu, j, C_[s2], xi) *
                                            GB(Gq, d, nO, j, A_[s2], ga, i, A_[s1], nu);
                                    acc[cidx(nO, s1, i, mu, nu) + (size_t)D * cidx(nO, s2, j, ga, xi)] +=
                                        w * (t1 + conj(t2));
                                }
                        }
}

/* Real-axis pair (eintshechi.F90:143-165): Gq = G(k+q, x+omega+i eta),
 * G = G(k, x+i eta').  acc += -i w (Gq_aa' - conj(Gq_a'a^T)) conj(G_cc'). */
static void chi0_real_pair(int nA, int nO, int d, int D, const zc *Gq, const zc *G, double w, zc *acc)
{
    static const int A_[4] = {1, 0, 1, 0}, C_[4] = {0, 0, 1, 1};
    for (int j = 0; j < nA; ++j)
        for (int i = 0; i < nA; ++i)
            for (int ga = 0; ga < nO; ++ga)
                for (int mu = 0; mu < nO; ++mu)
                    for (int xi = 0; xi < nO; ++xi)
                        for (int nu = 0; nu < nO; ++nu)
                            for (int s1 = 0; s1 < 4; ++s1)
                                for (int s2 = 0; s2 < 4; ++s2) {
                                    zc a = GB(Gq, d, nO, i, A_[s1], nu, j, A_[s2], ga) -
                                           conj(GB(Gq, d, nO, j, A_[s2], ga, i, A_[s1], nu));
                                    zc c = conj(GB(G, d, nO, i, C_[s1], mu, j, C_[s2], xi));
                                    acc[cidx(nO, s1, i, mu, nu) + (size_t)D * cidx(nO, s2, j, ga, xi)] +=
                                        -w * I * a * c;
                                }
}

/* chi0(q, omega) for one frequency (eintshechi.F90:2-178).
 *   imag list: n_im points (k = kim[3*p..], energy y = yim[p], weight wim[p]
 *              = wght(E) * w_k)
 *   real axis: used when |omega| >= 1e-15: nk_re k-points (kre, wre) times
 *              n_x energies (x2, p2) in [Ef-omega, Ef]
 * out = (sum_imag + sum_real)/(2 pi), D x D column-major.
 * Returns 0 or the first nonzero LU info. */
int or_chi0(const or_sys *s, const double *q, double omega, double Ef, double eta, double etap,
            long n_im, const double *kim, const double *yim, const double *wim,
            long nk_re, const double *kre, const double *wre, int n_x, const double *x2, const double *p2,
            zc *out, int nthreads)
{
    const int nA = s->nA, nO = s->nOrb, d = s->d, D = s->D;
    int status = 0;
    long n_re = (fabs(omega) >= 1e-15) ? (long)n_x * nk_re : 0;
    memset(out, 0, sizeof(zc) * (size_t)D * D);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        zc *acc = (zc *)calloc((size_t)D * D, sizeof(zc));
        zc *facc = (zc *)calloc((size_t)D * D, sizeof(zc));
        zc *Gq = (zc *)malloc(sizeof(zc) * d * d);
        zc *G = (zc *)malloc(sizeof(zc) * d * d);
        int st = 0;
#pragma omp for schedule(dynamic, 4) nowait
        for (long p = 0; p < n_im; ++p) {
            double k[3] = {kim[3 * p], kim[3 * p + 1], kim[3 * p + 2]};
            double kq[3] = {k[0] + q[0], k[1] + q[1], k[2] + q[2]};
            int r1 = or_green(s, Ef + omega, yim[p] + eta, kq, Gq);
            int r2 = or_green(s, Ef, yim[p] + etap, k, G);
            if ((r1 || r2) && !st) st = r1 ? r1 : r2;
            chi0_imag_pair(nA, nO, d, D, Gq, G, wim[p], acc);
        }
#pragma omp for schedule(dynamic, 4) nowait
        for (long p = 0; p < n_re; ++p) {
            long ie = p / nk_re, ik = p % nk_re;
            double k[3] = {kre[3 * ik], kre[3 * ik + 1], kre[3 * ik + 2]};
            double kq[3] = {k[0] + q[0], k[1] + q[1], k[2] + q[2]};
            int r1 = or_green(s, x2[ie] + omega, eta, kq, Gq);
            int r2 = or_green(s, x2[ie], etap, k, G);
            if ((r1 || r2) && !st) st = r1 ? r1 : r2;
            chi0_real_pair(nA, nO, d, D, Gq, G, p2[ie] * wre[ik], facc);
        }
#pragma omp critical
        {
            for (size_t i = 0; i < (size_t)D * D; ++i) out[i] += acc[i] + facc[i];
            if (st && !status) status = st;
        }
        free(acc); free(facc); free(Gq); free(G);
    }
    const double tpi = 2.0 * acos(-1.0);
    for (size_t i = 0; i < (size_t)D * D; ++i) out[i] /= tpi;
    return status;
}

/* ------------------------------------------------------------------------ */
/* Dyson step and orbital trace                                              */
/* ------------------------------------------------------------------------ */

/* chi = (I + chi0 U)^-1 chi0 with explicit inverse, as calculate_chi.F90:84-89.
 * Returns the LU info of (I + chi0 U). */
int or_dyson(int D, const zc *chi0, const zc *U, zc *chi)
{
    zc *T = (zc *)malloc(sizeof(zc) * (size_t)D * D);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < D; ++c) {
        for (int r = 0; r < D; ++r) T[r + (size_t)D * c] = (r == c) ? 1.0 : 0.0;
        for (int k = 0; k < D; ++k) {
            zc u = U[k + (size_t)D * c];
            if (u != 0.0)
                for (int r = 0; r < D; ++r) T[r + (size_t)D * c] += chi0[r + (size_t)D * k] * u;
        }
    }
    int info = or_zinv(D, T);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < D; ++c) {
        for (int r = 0; r < D; ++r) chi[r + (size_t)D * c] = 0.0;
        for (int k = 0; k < D; ++k) {
            zc b = chi0[k + (size_t)D * c];
            for (int r = 0; r < D; ++r) chi[r + (size_t)D * c] += T[r + (size_t)D * k] * b;
        }
    }
    free(T);
    return info;
}

/* schi[(sigma,i),(sigma',j)] = sum_{mu,nu} chi[(sigma,i,mu,mu),(sigma',j,nu,nu)]
 * with (sigma,i) -> sigma*nA + i (calculate_chi.F90:91-107, sigmai2i). */
void or_trace(int nA, int nOrb, const zc *chi, zc *schi)
{
    int D = 4 * nOrb * nOrb * nA, n4 = 4 * nA;
    memset(schi, 0, sizeof(zc) * n4 * n4);
    for (int j = 0; j < nA; ++j)
        for (int i = 0; i < nA; ++i)
            for (int sp = 0; sp < 4; ++sp)
                for (int sg = 0; sg < 4; ++sg)
                    for (int nu = 0; nu < nOrb; ++nu)
                        for (int mu = 0; mu < nOrb; ++mu)
                            schi[(sg * nA + i) + n4 * (sp * nA + j)] +=
                                chi[cidx(nOrb, sg, i, mu, mu) + (size_t)D * cidx(nOrb, sp, j, nu, nu)];
}

/* Index map exported for tests of the chi-basis ordering. */
int or_cidx(int nOrb, int sigma, int i, int mu, int nu) { return cidx(nOrb, sigma, i, mu, nu); }
