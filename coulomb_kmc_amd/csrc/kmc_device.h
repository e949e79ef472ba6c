// Device code shared by the library's translation units: recurrence tables, basis
// routines, kernel argument structs, the kernels templated on the expansion order L
// and their launchers.  kmc_amd.hip holds the kernels that do not depend on L and
// the host code; kmc_order.hip instantiates the L-templated launchers for one order
// per object, so the orders compile in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <string>
#include <vector>
#include <algorithm>
#include <unordered_map>

#include "../../include/kmc_amd.h"

// ============================================================================
// errors
// ============================================================================
static thread_local std::string g_err;
static int set_err(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return set_err(KMC_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));   \
    } while (0)

// ============================================================================
// compile-time recurrence tables
// ============================================================================
constexpr int KMIN_L = 2;
constexpr int KMAX_L = 20;

constexpr double c_sqrt(double x) {
    if (x <= 0.0) return 0.0;
    double g = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 200; i++) {
        const double n = 0.5 * (g + x / g);
        if (n == g) break;
        g = n;
    }
    return g;
}
struct RecTab {
    double a[KMAX_L + 1][KMAX_L + 1];   // alpha_lm = (2l-1)/sqrt(l^2-m^2)
    double b[KMAX_L + 1][KMAX_L + 1];   // beta_lm  = sqrt(((l-1)^2-m^2)/(l^2-m^2))
    double s[KMAX_L + 1];               // s_m = (-1)^m (2m-1)!!/sqrt((2m)!)
};
constexpr RecTab make_rectab() {
    RecTab t{};
    t.s[0] = 1.0;
    for (int m = 1; m <= KMAX_L; m++) t.s[m] = -t.s[m - 1] * c_sqrt((2.0 * m - 1.0) / (2.0 * m));
    for (int m = 0; m <= KMAX_L; m++)
        for (int l = m + 1; l <= KMAX_L; l++) {
            const double d = (double)(l * l - m * m);
            t.a[l][m] = (2.0 * l - 1.0) / c_sqrt(d);
            const double e = (double)((l - 1) * (l - 1) - m * m);
            t.b[l][m] = e > 0.0 ? c_sqrt(e / d) : 0.0;
        }
    return t;
}
static constexpr RecTab kRec = make_rectab();

__host__ __device__ constexpr int hidx(int l, int m) { return l * (l + 1) / 2 + m; }
__host__ __device__ constexpr int nhalf(int L) { return L * (L + 1) / 2; }

// ============================================================================
// small device helpers
// ============================================================================
struct Cplx {
    double re, im;
};
__device__ __forceinline__ Cplx cmul(Cplx a, Cplx b) {
    return {fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re)};
}
template <int M>
__device__ __forceinline__ Cplx cpow(Cplx u) {
    if constexpr (M == 0) {
        return {1.0, 0.0};
    } else if constexpr (M == 1) {
        return u;
    } else if constexpr (M % 2 == 0) {
        const Cplx h = cpow<M / 2>(u);
        return cmul(h, h);
    } else {
        return cmul(cpow<M - 1>(u), u);
    }
}

// 1/sqrt(x) to full FP64 accuracy: FP32 hardware estimate, two Newton steps in FP64.
__device__ __forceinline__ double rsqrt_fp64(double x) {
    double y = (double)__builtin_amdgcn_rsqf((float)x);
    double h = fma(-0.5 * x * y, y, 0.5);
    y = fma(y, h, y);
    h = fma(-0.5 * x * y, y, 0.5);
    return fma(y, h, y);
}
// 1/sqrt(x) with one Newton step (relative error ~3e-14): the mixed-precision path.
__device__ __forceinline__ double rsqrt_mixed(double x) {
    const double y = (double)__builtin_amdgcn_rsqf((float)x);
    const double h = fma(-0.5 * x * y, y, 0.5);
    return fma(y, h, y);
}
// plain FP64 (library sqrt and division, as the reference's 1.0/sqrt(r2))
__device__ __forceinline__ double rsqrt_plain(double x) { return 1.0 / sqrt(x); }

template <bool MIXED>
__device__ __forceinline__ double rsqrt_sel(double x) {
    if constexpr (MIXED) return rsqrt_mixed(x);
    else return rsqrt_plain(x);
}

struct Geom {
    double E;        // box side
    double w;        // leaf width
    double c0;       // centre of leaf 0 along an axis: -E/2 + w/2
    int ncps;        // leaves per side
    int bc;
};

// leaf of a coordinate: truncate (p + E/2)/w and clamp (coulomb_kmc/kmc_fmm.py:229-242)
__host__ __device__ inline int bin_axis(double p, double E, double w, int ncps) {
    long long c = (long long)((0.5 * E + p) / w);
    if (c < 0) c = 0;
    if (c > ncps - 1) c = ncps - 1;
    return (int)c;
}

// Compensated summation.  two_sum is Knuth's error-free transformation (s + e == a + b
// exactly, s = fl(a + b)); a CSum carries a running sum and the sum of the rounding
// errors of its additions, so a sum of k terms is good to ~1 ulp of the result plus
// k eps^2 of the terms instead of k eps of the partial sums.  (No -ffast-math
// anywhere in this build, so the compiler keeps these operations as written.)
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
struct CSum {
    double s, c;
    __device__ __forceinline__ CSum() : s(0.0), c(0.0) {}
    __device__ __forceinline__ void add(double v) {
        double e;
        two_sum(s, v, s, e);
        c += e;
    }
    // a * b added with the rounding error of the product (FMA two-product)
    __device__ __forceinline__ void add_prod(double a, double b) {
        const double p = a * b;
        const double pe = fma(a, b, -p);
        double e;
        two_sum(s, p, s, e);
        c += e + pe;
    }
    __device__ __forceinline__ void add(const CSum& o) {
        double e;
        two_sum(s, o.s, s, e);
        c += e + o.c;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};
// compensated sum over the lanes of groups of G consecutive lanes (G a power of two <= 64)
template <int G>
__device__ __forceinline__ CSum csum_group(CSum v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        CSum w;
        w.s = __shfl_xor(v.s, o, G);
        w.c = __shfl_xor(v.c, o, G);
        v.add(w);
    }
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// stencil offsets, x fastest (coulomb_kmc/common.py:51-81)
__device__ __forceinline__ void stencil(int k, int& ox, int& oy, int& oz) {
    ox = k % 3 - 1;
    oy = (k / 3) % 3 - 1;
    oz = k / 9 - 1;
}

// ============================================================================
// accept sweep / initial P2L: every leaf receives q I_l^m conj(...) from sources
// ============================================================================
// Sources come either from the accepted move (old position with -q, new position
// with +q) or from a range of particles (+q, initialisation).  For '27'/PBC each
// source is applied with its 27 images (image shift o*E, cell shift o*ncps) and the
// reference's well-separated test (coulomb_kmc/kmc_octal.py:966-975).
struct MoveRec {
    double old_pos[3];
    double new_pos[3];
    double q;
    double dE;
    int32_t pid;
    int32_t old_cell;
    int32_t new_cell;
    int32_t valid;
};

struct SweepArgs {
    Geom g;
    double2* lexp;             // [ncells][nh]
    int ncells;
    int mode;                  // 0: move record, 1: particle range
    const MoveRec* rec;
    const double4* src_pos4;   // mode 1
    const int32_t* src_cell;   // mode 1
    int src_begin, src_end;
};

// Rescaled column recurrence of the sweep: e_l = c_l / Gam_l obeys
//   e_l = a2_lm zr e_{l-1} - ir2 e_{l-2},  Gam_l = beta_lm Gam_{l-2},  a2_lm = alpha_lm Gam_{l-1} / Gam_l,
// one multiply less per coefficient; accumulators are multiplied by Gam_l s_m (gs, the
// column's diagonal factor s_m folded in) once per leaf.
struct RecTab2 {
    double a2[KMAX_L + 1][KMAX_L + 1];
    double gam[KMAX_L + 1][KMAX_L + 1];
    double gs[KMAX_L + 1][KMAX_L + 1];
};
constexpr RecTab2 make_rectab2() {
    RecTab2 t{};
    for (int m = 0; m <= KMAX_L; m++) {
        t.gam[m][m] = 1.0;
        if (m + 1 <= KMAX_L) {
            t.gam[m + 1][m] = 1.0;
            t.a2[m + 1][m] = kRec.a[m + 1][m];
        }
        for (int l = m + 2; l <= KMAX_L; l++) {
            t.gam[l][m] = kRec.b[l][m] * t.gam[l - 2][m];
            t.a2[l][m] = kRec.a[l][m] * t.gam[l - 1][m] / t.gam[l][m];
        }
        for (int l = m; l <= KMAX_L; l++) t.gs[l][m] = t.gam[l][m] * kRec.s[m];
    }
    return t;
}
static constexpr RecTab2 kRec2 = make_rectab2();

// one column m of NI source-images: acc[k] += e_{m+k} * W_j, k = 0 .. L-m-1, j = 0 .. NI-1.
// The NI recurrences are independent dependency chains (instruction-level
// parallelism); each accumulator still receives the images in order j = 0, 1, ...
// orders below LC only (LC <= L: the blocked sweep's truncation buckets)
template <int L, int M, int NI, int LC = L>
__device__ __forceinline__ void sweep_column(const double* zr, const double* ir2, const Cplx* W, double* ar,
                                             double* ai) {
    constexpr int LT = LC < L ? LC : L;
    double e2[NI], e1[NI];
#pragma unroll
    for (int j = 0; j < NI; j++) {
        e2[j] = 0.0;
        e1[j] = 1.0;
        ar[0] += W[j].re;
        ai[0] += W[j].im;
    }
#pragma unroll
    for (int k = 1; k < LT - M; k++) {
        const int l = M + k;
#pragma unroll
        for (int j = 0; j < NI; j++) {
            const double e = fma(kRec2.a2[l][M] * zr[j], e1[j], -ir2[j] * e2[j]);
            e2[j] = e1[j];
            e1[j] = e;
        }
#pragma unroll
        for (int j = 0; j < NI; j++) {
            ar[k] = fma(e1[j], W[j].re, ar[k]);
            ai[k] = fma(e1[j], W[j].im, ai[k]);
        }
    }
}

template <int L, int M>
__device__ __forceinline__ void sweep_store_column(double2* le, const double* ar, const double* ai) {
#pragma unroll
    for (int k = 0; k < L - M; k++) {
        const double gm = kRec2.gs[M + k][M];
        double2 v = le[hidx(M + k, M)];
        v.x = fma(gm, ar[k], v.x);
        v.y = fma(gm, ai[k], v.y);
        le[hidx(M + k, M)] = v;
    }
}

// columns (P, L-1-P) of pair P, NI images, into accumulator row ar/ai (first L-P entries: column P)
// column M with weight U = base conj((x+iy)/r^2)^M into the accumulator row (its factor
// s_M is applied with Gam when the sums are stored)
template <int L, int M, int NI, int LC = L>
__device__ __forceinline__ void sweep_weighted_column(const double* zr, const double* ir2, const Cplx* U,
                                                      double* ar, double* ai) {
    if constexpr (M < LC) sweep_column<L, M, NI, LC>(zr, ir2, U, ar, ai);
}
// low columns P0 + J, J = 0 .. NPE-1 (ascending powers: one multiply per column)
template <int L, int P0, int NPE, int NI, int LC, int J = 0>
__device__ __forceinline__ void sweep_low_cols(const double* zr, const double* ir2, const Cplx* u, Cplx* U,
                                               const double* base, double (*ar)[L + 1], double (*ai)[L + 1]) {
    if constexpr (J < NPE) {
        if constexpr (J > 0) {
#pragma unroll
            for (int j = 0; j < NI; j++) U[j] = cmul(U[j], u[j]);
        }
        sweep_weighted_column<L, P0 + J, NI, LC>(zr, ir2, U, ar[J], ai[J]);
        sweep_low_cols<L, P0, NPE, NI, LC, J + 1>(zr, ir2, u, U, base, ar, ai);
    }
}
// high columns L-1-P0-J for J = NPE-1 down to 0 (ascending powers), skipping the
// middle column that equals its low partner when L is odd
template <int L, int P0, int NPE, int NI, int LC, int J>
__device__ __forceinline__ void sweep_high_cols(const double* zr, const double* ir2, const Cplx* u, Cplx* U,
                                                const double* base, double (*ar)[L + 1], double (*ai)[L + 1]) {
    if constexpr (J >= 0) {
        constexpr int M1 = P0 + J, M2 = L - 1 - P0 - J;
        if constexpr (M2 != M1)
            sweep_weighted_column<L, M2, NI, LC>(zr, ir2, U, ar[J] + (L - M1), ai[J] + (L - M1));
        if constexpr (J > 0) {
#pragma unroll
            for (int j = 0; j < NI; j++) U[j] = cmul(U[j], u[j]);
        }
        sweep_high_cols<L, P0, NPE, NI, LC, J - 1>(zr, ir2, u, U, base, ar, ai);
    }
}
// the column pairs (P, L-1-P), P = P0 .. P0+NP-1, of one lane: the powers of u run up
// through the low columns and on (one cpow) through the high ones
// LC: orders l >= LC are left out (the blocked sweep's truncation buckets; L: none).
// Columns ascend within the low and within the high part, so a column at or above LC
// ends its part.
// U0[j] = base u^P0, the weight of column P0 (the power chain carries base along)
template <int L, int P0, int NP, int NI, int LC = L>
__device__ __forceinline__ void sweep_group_cols_from(const double* zr, const double* ir2, const Cplx* u,
                                                      const Cplx* U0, const double* base, double (*ar)[L + 1],
                                                      double (*ai)[L + 1]) {
    constexpr int NPAIR = (L + 1) / 2;
    constexpr int NPE = (P0 + NP <= NPAIR) ? NP : NPAIR - P0;   // pairs in this group
    if constexpr (NPE > 0 && P0 < LC) {
        Cplx U[NI];
#pragma unroll
        for (int j = 0; j < NI; j++) U[j] = U0[j];
        sweep_low_cols<L, P0, NPE, NI, LC>(zr, ir2, u, U, base, ar, ai);
        constexpr int LOW_TOP = P0 + NPE - 1;          // power held in U now
        constexpr int HIGH_BOTTOM = L - 1 - LOW_TOP;    // lowest high column
        if constexpr (HIGH_BOTTOM > LOW_TOP && HIGH_BOTTOM < LC) {
#pragma unroll
            for (int j = 0; j < NI; j++) U[j] = cmul(U[j], cpow<HIGH_BOTTOM - LOW_TOP>(u[j]));
            sweep_high_cols<L, P0, NPE, NI, LC, NPE - 1>(zr, ir2, u, U, base, ar, ai);
        } else if constexpr (HIGH_BOTTOM == LOW_TOP && NPE > 1 && HIGH_BOTTOM + 1 < LC) {
            // L odd: the last low column is the middle one; the high columns start one above
#pragma unroll
            for (int j = 0; j < NI; j++) U[j] = cmul(U[j], u[j]);
            sweep_high_cols<L, P0, NPE, NI, LC, NPE - 2>(zr, ir2, u, U, base, ar, ai);
        }
    }
}
template <int L, int P0, int NP, int NI, int LC = L>
__device__ __forceinline__ void sweep_group_cols(const double* zr, const double* ir2, const Cplx* u,
                                                 const double* base, double (*ar)[L + 1], double (*ai)[L + 1]) {
    constexpr int NPAIR = (L + 1) / 2;
    if constexpr (P0 < NPAIR && P0 < LC) {
        Cplx U0[NI];
#pragma unroll
        for (int j = 0; j < NI; j++) {
            const Cplx c = cpow<P0>(u[j]);
            U0[j] = Cplx{base[j] * c.re, base[j] * c.im};
        }
        sweep_group_cols_from<L, P0, NP, NI, LC>(zr, ir2, u, U0, base, ar, ai);
    }
}
template <int L, int P0, int NP, int J = 0>
__device__ __forceinline__ void sweep_group_store(double2* le, double (*ar)[L + 1], double (*ai)[L + 1]) {
    if constexpr (J < NP && P0 + J < (L + 1) / 2) {
        constexpr int M1 = P0 + J, M2 = L - 1 - M1;
        sweep_store_column<L, M1>(le, ar[J], ai[J]);
        if constexpr (M2 != M1) sweep_store_column<L, M2>(le, ar[J] + (L - M1), ai[J] + (L - M1));
        sweep_group_store<L, P0, NP, J + 1>(le, ar, ai);
    }
}

// NI source images (displacement leaf centre -> image, charge or 0 when not well separated)
template <int L, int P0, int NP, int NI>
__device__ __forceinline__ void sweep_images(const double* dx, double dy, double dz, const double* qe,
                                             double (*ar)[L + 1], double (*ai)[L + 1]) {
    double zr[NI], ir2[NI], base[NI];
    Cplx u[NI];
#pragma unroll
    for (int j = 0; j < NI; j++) {
        double r2 = fma(dx[j], dx[j], fma(dy, dy, dz * dz));
        r2 = qe[j] != 0.0 ? r2 : 1.0;
        const double ir = rsqrt_fp64(r2);
        ir2[j] = ir * ir;
        zr[j] = dz * ir2[j];
        u[j] = Cplx{dx[j] * ir2[j], -dy * ir2[j]};     // conj((x+iy)/r^2)
        base[j] = qe[j] * ir;
    }
    sweep_group_cols<L, P0, NP, NI>(zr, ir2, u, base, ar, ai);
}
template <int L, int P0, int NP, int NI, int LC = L>
__device__ __forceinline__ void sweep_images_xyz(const double* dx, const double* dy, const double* dz,
                                                 const double* qe, double (*ar)[L + 1], double (*ai)[L + 1]) {
    double zr[NI], ir2[NI], base[NI];
    Cplx u[NI];
#pragma unroll
    for (int j = 0; j < NI; j++) {
        double r2 = fma(dx[j], dx[j], fma(dy[j], dy[j], dz[j] * dz[j]));
        r2 = qe[j] != 0.0 ? r2 : 1.0;
        const double ir = rsqrt_fp64(r2);
        ir2[j] = ir * ir;
        zr[j] = dz[j] * ir2[j];
        u[j] = Cplx{dx[j] * ir2[j], -dy[j] * ir2[j]};
        base[j] = qe[j] * ir;
    }
    sweep_group_cols<L, P0, NP, NI, LC>(zr, ir2, u, base, ar, ai);
}

__device__ __forceinline__ bool near1(int d) { return d < 2 && d > -2; }

// pairs P = G*NP .. G*NP+NP-1 of columns (P, L-1-P) for one leaf (one lane).
// PER ('27'/PBC): the 27 images of a source in stencil order (x fastest), of which
// this lane takes the z-layers oz = z0 .. z1 (all three unless the images are split
// over the waves of the workgroup); free space: the source itself.
template <int L, int G, int NP, bool PER>
__device__ __forceinline__ void sweep_group(const SweepArgs& a, int cell, int z0, int z1, double (*ar)[L + 1],
                                            double (*ai)[L + 1]) {
    constexpr int P0 = G * NP;
    const int n = a.g.ncps;
    const double E = a.g.E;
    const int tx = cell % n, ty = (cell / n) % n, tz = cell / (n * n);
    const double cx = a.g.c0 + tx * a.g.w;
    const double cy = a.g.c0 + ty * a.g.w;
    const double cz = a.g.c0 + tz * a.g.w;
    int nsrc;
    if (a.mode == 0) nsrc = a.rec->valid ? 2 : 0;
    else nsrc = a.src_end - a.src_begin;

    if constexpr (PER && NP == 2) {
        if (a.mode == 0 && nsrc == 2) {
            // the move: each image of the old (-q) and the new (+q) position together
            // (two independent recurrences per column; with one column pair per lane the
            // extra registers would spill)
            const MoveRec* rc = a.rec;
            double dx0[2], dy0[2], dz0[2], sq[2];
            int dcx[2], dcy[2], dcz[2];
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const double* p = s == 0 ? rc->old_pos : rc->new_pos;
                const int sc = s == 0 ? rc->old_cell : rc->new_cell;
                sq[s] = s == 0 ? -1.0 * rc->q : rc->q;
                dx0[s] = p[0] - cx; dy0[s] = p[1] - cy; dz0[s] = p[2] - cz;
                dcx[s] = sc % n - tx; dcy[s] = (sc / n) % n - ty; dcz[s] = sc / (n * n) - tz;
            }
#pragma unroll 1
            for (int oz = z0; oz <= z1; oz++) {
#pragma unroll 1
                for (int oy = -1; oy <= 1; oy++) {
#pragma unroll 1
                    for (int ox = -1; ox <= 1; ox++) {
                        double dx[2], dy[2], dz[2], qe[2];
#pragma unroll
                        for (int s = 0; s < 2; s++) {
                            dx[s] = fma((double)ox, E, dx0[s]);
                            dy[s] = fma((double)oy, E, dy0[s]);
                            dz[s] = fma((double)oz, E, dz0[s]);
                            const bool nr = near1(dcx[s] + ox * n) && near1(dcy[s] + oy * n) && near1(dcz[s] + oz * n);
                            qe[s] = nr ? 0.0 : sq[s];
                        }
                        sweep_images_xyz<L, P0, NP, 2>(dx, dy, dz, qe, ar, ai);
                    }
                }
            }
            return;
        }
    }
    for (int s = 0; s < nsrc; s++) {
        double sx, sy, sz, sq;
        int sc;
        if (a.mode == 0) {
            const double* p = s == 0 ? a.rec->old_pos : a.rec->new_pos;
            sx = p[0]; sy = p[1]; sz = p[2];
            sq = s == 0 ? -1.0 * a.rec->q : a.rec->q;
            sc = s == 0 ? a.rec->old_cell : a.rec->new_cell;
        } else {
            const double4 p = a.src_pos4[a.src_begin + s];
            sx = p.x; sy = p.y; sz = p.z; sq = p.w;
            sc = a.src_cell[a.src_begin + s];
        }
        const int dcx = sc % n - tx, dcy = (sc / n) % n - ty, dcz = sc / (n * n) - tz;
        const double dx0 = sx - cx, dy0 = sy - cy, dz0 = sz - cz;
        if constexpr (!PER) {
            const double dx[1] = {dx0};
            const double qe[1] = {(near1(dcx) && near1(dcy) && near1(dcz)) ? 0.0 : sq};
            sweep_images<L, P0, NP, 1>(dx, dy0, dz0, qe, ar, ai);
        } else {
            const double dxm = dx0 - E, dxp = dx0 + E;
            const bool nxm = near1(dcx - n), nx0 = near1(dcx), nxp = near1(dcx + n);
#pragma unroll 1
            for (int oz = z0; oz <= z1; oz++) {
                const double dz = fma((double)oz, E, dz0);
                const bool nz = near1(dcz + oz * n);
#pragma unroll 1
                for (int oy = -1; oy <= 1; oy++) {
                    const double dy = fma((double)oy, E, dy0);
                    const bool nyz = nz && near1(dcy + oy * n);
#pragma unroll 1
                    for (int ox = -1; ox <= 1; ox++) {
                        const double dx[1] = {ox < 0 ? dxm : (ox == 0 ? dx0 : dxp)};
                        const bool nxs = ox < 0 ? nxm : (ox == 0 ? nx0 : nxp);
                        const double qe[1] = {(nyz && nxs) ? 0.0 : sq};
                        sweep_images<L, P0, NP, 1>(dx, dy, dz, qe, ar, ai);
                    }
                }
            }
        }
    }
}

template <int L, int NP, bool PER, int G0>
__device__ __forceinline__ void sweep_dispatch(const SweepArgs& a, int cell, int group, int z0, int z1,
                                               double (*ar)[L + 1], double (*ai)[L + 1]) {
    constexpr int NG = ((L + 1) / 2 + NP - 1) / NP;
    if constexpr (G0 < NG) {
        if (group == G0) sweep_group<L, G0, NP, PER>(a, cell, z0, z1, ar, ai);
        else sweep_dispatch<L, NP, PER, G0 + 1>(a, cell, group, z0, z1, ar, ai);
    }
}
template <int L, int NP, int G0>
__device__ __forceinline__ void sweep_store_dispatch(double2* le, int group, double (*ar)[L + 1], double (*ai)[L + 1]) {
    constexpr int NG = ((L + 1) / 2 + NP - 1) / NP;
    if constexpr (G0 < NG) {
        if (group == G0) sweep_group_store<L, G0 * NP, NP>(le, ar, ai);
        else sweep_store_dispatch<L, NP, G0 + 1>(le, group, ar, ai);
    }
}

// scale a group's sums in place by Gam (rescaled recurrence), then copy them to a leaf row
template <int L, int M>
__device__ __forceinline__ void sweep_scale_column(double* ar, double* ai) {
#pragma unroll
    for (int k = 0; k < L - M; k++) {
        ar[k] *= kRec2.gs[M + k][M];
        ai[k] *= kRec2.gs[M + k][M];
    }
}
template <int L, int M>
__device__ __forceinline__ void sweep_stage_column(double2* row, const double* ar, const double* ai) {
#pragma unroll
    for (int k = 0; k < L - M; k++) row[hidx(M + k, M)] = make_double2(ar[k], ai[k]);
}
template <int L, int P0, int NP, bool SCALE, int J = 0>
__device__ __forceinline__ void sweep_group_stage(double2* row, double (*ar)[L + 1], double (*ai)[L + 1]) {
    if constexpr (J < NP && P0 + J < (L + 1) / 2) {
        constexpr int M1 = P0 + J, M2 = L - 1 - M1;
        if constexpr (SCALE) sweep_scale_column<L, M1>(ar[J], ai[J]);
        else sweep_stage_column<L, M1>(row, ar[J], ai[J]);
        if constexpr (M2 != M1) {
            if constexpr (SCALE) sweep_scale_column<L, M2>(ar[J] + (L - M1), ai[J] + (L - M1));
            else sweep_stage_column<L, M2>(row, ar[J] + (L - M1), ai[J] + (L - M1));
        }
        sweep_group_stage<L, P0, NP, SCALE, J + 1>(row, ar, ai);
    }
}
template <int L, int NP, bool SCALE, int G0>
__device__ __forceinline__ void sweep_stage_dispatch(double2* row, int group, double (*ar)[L + 1], double (*ai)[L + 1]) {
    constexpr int NG = ((L + 1) / 2 + NP - 1) / NP;
    if constexpr (G0 < NG) {
        if (group == G0) sweep_group_stage<L, G0 * NP, NP, SCALE>(row, ar, ai);
        else sweep_stage_dispatch<L, NP, SCALE, G0 + 1>(row, group, ar, ai);
    }
}

// Lane = leaf (threadIdx.x), column-pair group = blockIdx.y.  With SPLIT = 3 the
// workgroup is 64 leaves x 3 waves and wave threadIdx.y takes the image layer
// oz = threadIdx.y - 1; the two upper waves hand their sums to wave 0 through LDS
// (fixed order), which alone updates the expansions.
template <int L, int NP, bool PER, int SPLIT>
__global__ void __launch_bounds__(256) k_sweep(SweepArgs a) {
    constexpr int W = L + 1;                 // accumulator entries per pair
    const int cell0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = cell0 < a.ncells;
    const int cell = live ? cell0 : a.ncells - 1;
    const int group = (int)blockIdx.y;
    double ar[NP][W], ai[NP][W];
#pragma unroll
    for (int j = 0; j < NP; j++)
#pragma unroll
        for (int k = 0; k < W; k++) { ar[j][k] = 0.0; ai[j][k] = 0.0; }
    if constexpr (SPLIT == 1) {
        sweep_dispatch<L, NP, PER, 0>(a, cell, group, -1, 1, ar, ai);
    } else {
        static_assert(SPLIT == 3 && PER, "image split needs the 27-image stencil");
        __shared__ double red[2][NP][2 * W][64];
        const int sl = (int)threadIdx.y;
        sweep_dispatch<L, NP, PER, 0>(a, cell, group, sl - 1, sl - 1, ar, ai);
        if (sl > 0) {
#pragma unroll
            for (int j = 0; j < NP; j++)
#pragma unroll
                for (int k = 0; k < W; k++) {
                    red[sl - 1][j][2 * k][threadIdx.x] = ar[j][k];
                    red[sl - 1][j][2 * k + 1][threadIdx.x] = ai[j][k];
                }
        }
        __syncthreads();
        if (sl > 0) return;
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int j = 0; j < NP; j++)
#pragma unroll
                for (int k = 0; k < W; k++) {
                    ar[j][k] += red[u][j][2 * k][threadIdx.x];
                    ai[j][k] += red[u][j][2 * k + 1][threadIdx.x];
                }
    }
    if (!live) return;
    sweep_store_dispatch<L, NP, 0>(a.lexp + (size_t)cell * nhalf(L), group, ar, ai);
}

// Fused layout: a workgroup owns 64 consecutive leaves and all their column groups
// (wave threadIdx.y = group), so the update of the leaves' expansions is one
// contiguous block of 64 * nh coefficients.  The sums are staged through LDS in two
// halves of 32 leaves and added to the expansions with coalesced loads and stores
// (a lane's own columns are scattered over its leaf: written directly they touch a
// cache line per lane and the groups of a leaf would re-fetch the same lines).
constexpr int stage_stride(int nh) { return nh + ((9 - nh % 8) % 8); }   // = 1 (mod 8): conflict-free b128 rows
template <int L, int NP, bool PER>
__device__ __forceinline__ void sweep_fused_body(const SweepArgs& a) {
    constexpr int W = L + 1;
    constexpr int NH = nhalf(L);
    constexpr int ST = stage_stride(NH);
    __shared__ double2 stage[32 * ST];
    const int lane = threadIdx.x;
    const int group = __builtin_amdgcn_readfirstlane((int)threadIdx.y);   // one group per wave
    const int cell0 = blockIdx.x * 64;
    const int cell = cell0 + lane;
    double ar[NP][W], ai[NP][W];
#pragma unroll
    for (int j = 0; j < NP; j++)
#pragma unroll
        for (int k = 0; k < W; k++) { ar[j][k] = 0.0; ai[j][k] = 0.0; }
    sweep_dispatch<L, NP, PER, 0>(a, cell, group, -1, 1, ar, ai);
    sweep_stage_dispatch<L, NP, true, 0>(nullptr, group, ar, ai);
    const int tid = threadIdx.x + 64 * threadIdx.y;
    const int nthr = 64 * blockDim.y;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
        __syncthreads();
        if ((lane >> 5) == half) sweep_stage_dispatch<L, NP, false, 0>(stage + (lane & 31) * ST, group, ar, ai);
        __syncthreads();
        double2* base = a.lexp + (size_t)(cell0 + 32 * half) * NH;
        // eight loads in flight per thread before their adds and stores
        constexpr int U = 8;
#pragma unroll 1
        for (int e0 = tid; e0 < 32 * NH; e0 += nthr * U) {
            double2 v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * nthr;
                if (e < 32 * NH) v[u] = base[e];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * nthr;
                if (e < 32 * NH) {
                    const int j = e / NH, h = e - j * NH;
                    const double2 sv = stage[j * ST + h];
                    v[u].x += sv.x;
                    v[u].y += sv.y;
                    base[e] = v[u];
                }
            }
        }
    }
}
// ---------------------------------------------------------------------------
// Accept sweep over 4x4x4 blocks of leaves, orders truncated where they cannot matter.
//
// The hop's P2L into a leaf (centre c) from a source image of charge q at s changes
// the potential the leaf's expansion gives anywhere in the leaf (|x - c| <= rho, the
// leaf's half diagonal) by the sum over orders l of q |x-c|^l / |s-c|^{l+1} P_l(cos)
// (addition theorem; |P_l| <= 1).  The orders l >= lc together change it by at most
//     |q| (rho/D)^lc / (D - rho),   D = min |s - c| over the block's leaf centres,
// so they are skipped when that bound is below SWEEP_TAU |q| (2^-60 per unit charge,
// ~1e-18 for potentials of order one): the leaf's expansion then differs from the
// full P2L by less than the rounding of its own lowest coefficients, and dE of any
// later proposal by less than 54 SWEEP_TAU |q| per accept.  lc is the same for every
// leaf of the block, so the truncation is uniform over each wave: a workgroup owns a
// 4x4x4 block of leaves (lane = leaf, wave = column-pair group, as in the fused
// layout), whose 64 expansions are 16 runs of 4 consecutive leaves for the staged
// update.  Images far from the whole block (most of the 27 of each source in a
// periodic box) keep only their first ~9-12 orders of 16.
// ---------------------------------------------------------------------------
constexpr double SWEEP_TAU = 8.673617379884035e-19;   // 2^-60

// smallest lc with |q| (rho/D)^lc / (D - rho) <= SWEEP_TAU |q|, capped at L; every
// order when D <= 3 rho.  A leaf whose 27-leaf stencil holds the image's leaf has its
// centre within 1.5 leaf widths of the image on each axis, so D < 1.5 sqrt(3) w = 3 rho:
// an image truncated below L has no such leaf in the block, and the truncated buckets
// need no near-leaf test.  (For L <= 20 the bound alone already keeps every order
// below 3 rho unless the leaves are ~1e12 units wide: (1/3)^L / (D - rho) > 2^-60.)
__device__ __forceinline__ int sweep_trunc_order(double D, double rho, int L) {
    if (!(D > 3.0 * rho)) return L;
    const double t = rho / D;
    double bound = 1.0 / (D - rho);
    int lc = 0;
    while (lc < L && bound > SWEEP_TAU) {
        bound *= t;
        lc++;
    }
    return lc;
}

// truncation buckets of the blocked sweep: orders below 8, 9, 10, 12, then all L.  In a
// periodic box at 2M charges on 64^3 leaves, L = 16, 90% of the block-image pairs have
// truncation orders 7-9; these buckets round the coefficient count up by 3% (8, 10, 12,
// 14, L: 8%).  More buckets cut the rounding further but make the kernel larger and
// slower (eight buckets: 0.1% rounding, 7% slower launches).
constexpr int SWEEP_NBUCKET = 5;
__host__ __device__ constexpr int sweep_bucket(int i, int L) {
    constexpr int B[SWEEP_NBUCKET - 1] = {8, 9, 10, 12};
    return i < SWEEP_NBUCKET - 1 ? (B[i] < L ? B[i] : L) : L;
}
// bucket index of an image of truncation order lc >= 1
__host__ __device__ constexpr int sweep_bucket_index(int lc, int L) {
    for (int i = 0; i < SWEEP_NBUCKET; i++)
        if (lc <= sweep_bucket(i, L)) return i;
    return SWEEP_NBUCKET - 1;
}
// orders evaluated for an image of truncation order lc: its bucket (0 for lc = 0, an
// image the blocked sweep skips)
__host__ __device__ constexpr int sweep_bucket_of(int lc, int L) {
    return lc <= 0 ? 0 : sweep_bucket(sweep_bucket_index(lc, L), L);
}
// per-block image lists of the blocked sweep (shared memory): bucket I's images at
// [I * SWEEP_LIST ..) in stencil order, each entry k | lc << 8, and the count of bucket
// I at [SWEEP_NBUCKET * SWEEP_LIST + I]
constexpr int SWEEP_LIST = 32;   // >= 27 images

// Column pairs of wave G in the blocked sweep: pair G and pair NPAIR-1-G (one pair when
// they coincide).  With every order kept each wave has 2(L+1) coefficients; with the
// orders truncated at LC the low columns of the low pairs stay and the high ones go, so
// the pairing from both ends keeps the waves' shares within a few coefficients of each
// other (consecutive pairs would leave the waves of high columns idle for far images).
template <int L, int G>
struct BlockPairs {
    static constexpr int NPAIR = (L + 1) / 2;
    static constexpr int A = G;
    static constexpr int B = NPAIR - 1 - G;   // == A: one pair
};
// one source image for the lane's leaf into wave G's accumulators (row 0: pair A, row 1: pair B)
// (NEAR: the image may be a near neighbour of the leaf, which then takes qe = 0)
template <int L, int G, int LC, bool NEAR>
__device__ __forceinline__ void block_image(double dx, double dy, double dz, double qe, double (*ar)[L + 1],
                                            double (*ai)[L + 1]) {
    using PP = BlockPairs<L, G>;
    double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    if constexpr (NEAR) r2 = qe != 0.0 ? r2 : 1.0;
    const double ir = rsqrt_fp64(r2);
    const double ir2 = ir * ir;
    const double zr = dz * ir2;
    const Cplx u = {dx * ir2, -dy * ir2};     // conj((x+iy)/r^2)
    const double base = qe * ir;
    sweep_group_cols<L, PP::A, 1, 1, LC>(&zr, &ir2, &u, &base, ar, ai);
    if constexpr (PP::B > PP::A) sweep_group_cols<L, PP::B, 1, 1, LC>(&zr, &ir2, &u, &base, ar + 1, ai + 1);
}

// Truncated images in wave pairs.  Waves G and G^1 (pairs G, NPAIR-1-G and G^1,
// NPAIR-1-(G^1)) share the four column pairs P0, P0+1, Q0, Q0+1 (P0 = G & ~1,
// Q0 = NPAIR-2-P0); for the images of the truncated buckets wave G evaluates all four
// for source G & 1 only (old position: even G, new: odd), so each source image's
// displacement, 1/r and weights are worked out by one wave of the pair instead of both.
// The two waves then swap the sums of each other's pairs through LDS.  The waves pair
// only when every wave of the workgroup has a partner with two distinct column pairs
// (NPAIR a multiple of 4: L = 15, 16 among the orders with truncated buckets): a wave
// left alone would still evaluate both sources and hold the workgroup at the swap
// (L = 12, three waves: the same sweep time, 5% slower 100k-charge steps).
template <int L, int G>
__host__ __device__ constexpr bool block_paired() {
    constexpr int NPAIR = (L + 1) / 2;
    return L > sweep_bucket(0, L) && NPAIR % 4 == 0 && G < NPAIR / 2;
}
// entries of pair P's two columns below order LC (the swap's message for that pair)
template <int L, int P, int LC>
__host__ __device__ constexpr int pair_entries() {
    constexpr int M1 = P, M2 = L - 1 - P;
    return (M1 < LC ? LC - M1 : 0) + ((M2 != M1 && M2 < LC) ? LC - M2 : 0);
}
// largest truncated bucket (orders below it; 0: none below L)
template <int L>
__host__ __device__ constexpr int sweep_top_trunc() {
    int t = 0;
    for (int i = 0; i < SWEEP_NBUCKET; i++)
        if (sweep_bucket(i, L) < L) t = sweep_bucket(i, L);
    return t;
}
// bucket index of the first bucket holding every order
template <int L>
__host__ __device__ constexpr int sweep_full_bucket() {
    return sweep_bucket_index(L, L);
}
// one source image into the four column pairs of wave pair P0 (rows P0, P0+1, Q0, Q0+1)
template <int L, int P0, int LC>
__device__ __forceinline__ void block_image_quad(double dx, double dy, double dz, double qe,
                                                 double (*ar)[L + 1], double (*ai)[L + 1]) {
    constexpr int Q0 = (L + 1) / 2 - 2 - P0;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double ir = rsqrt_fp64(r2);
    const double ir2 = ir * ir;
    const double zr = dz * ir2;
    const Cplx u = {dx * ir2, -dy * ir2};     // conj((x+iy)/r^2)
    const double base = qe * ir;
    sweep_group_cols<L, P0, 2, 1, LC>(&zr, &ir2, &u, &base, ar, ai);
    sweep_group_cols<L, Q0, 2, 1, LC>(&zr, &ir2, &u, &base, ar + 2, ai + 2);
}
// the truncated buckets' images of one source (in bucket, then stencil order) for a
// paired wave: all four pairs of the wave pair, rows P0, P0+1, Q0, Q0+1
template <int L, int G, bool PER, int I>
__device__ __forceinline__ void sweep_block_trunc(const SweepArgs& a, double dx0, double dy0, double dz0,
                                                  double q, const int* lists, double (*ar)[L + 1],
                                                  double (*ai)[L + 1]) {
    constexpr int LC = sweep_bucket(I, L);
    if constexpr (I < SWEEP_NBUCKET && LC < L) {
        constexpr int P0 = G & ~1;
        if constexpr (P0 < LC) {
            constexpr int R = PER ? 1 : 0;
            const double E = a.g.E;
            const int cnt = __builtin_amdgcn_readfirstlane(lists[SWEEP_NBUCKET * SWEEP_LIST + I]);
#pragma unroll 1
            for (int i = 0; i < cnt; i++) {
                const int e = __builtin_amdgcn_readfirstlane(lists[I * SWEEP_LIST + i]);
                const int k = e & 255, lc = e >> 8;
                if (lc <= P0) continue;   // (the image does not reach the pair's columns, the lowest is P0)
                const int ox = k % (2 * R + 1) - R, oy = (k / (2 * R + 1)) % (2 * R + 1) - R,
                          oz = k / ((2 * R + 1) * (2 * R + 1)) - R;
                block_image_quad<L, P0, LC>(fma((double)ox, E, dx0), fma((double)oy, E, dy0),
                                            fma((double)oz, E, dz0), q, ar, ai);
            }
        }
        sweep_block_trunc<L, G, PER, I + 1>(a, dx0, dy0, dz0, q, lists, ar, ai);
    }
}
// one pair's entries below order LC between a row and an LDS slot: PUT copies the row
// out, otherwise the slot is added to the row
template <int L, int P, int LC, bool PUT>
__device__ __forceinline__ void swap_pair(double2* slot, double* ar, double* ai) {
    constexpr int M1 = P, M2 = L - 1 - P;
    constexpr int N1 = M1 < LC ? LC - M1 : 0;
    constexpr int N2 = (M2 != M1 && M2 < LC) ? LC - M2 : 0;
#pragma unroll
    for (int k = 0; k < N1; k++) {
        if constexpr (PUT) {
            slot[k] = make_double2(ar[k], ai[k]);
        } else {
            const double2 v = slot[k];
            ar[k] += v.x;
            ai[k] += v.y;
        }
    }
#pragma unroll
    for (int k = 0; k < N2; k++) {
        if constexpr (PUT) {
            slot[N1 + k] = make_double2(ar[L - M1 + k], ai[L - M1 + k]);
        } else {
            const double2 v = slot[N1 + k];
            ar[L - M1 + k] += v.x;
            ai[L - M1 + k] += v.y;
        }
    }
}
// LDS slot size of the swap (entries of the larger message, conflict-free b128 rows)
template <int L>
__host__ __device__ constexpr int swap_stride() {
    constexpr int LT = sweep_top_trunc<L>();
    int m = 0;
    for (int p = 0; p < (L + 1) / 2; p++) {
        const int m1 = p, m2 = L - 1 - p;
        const int n = (m1 < LT ? LT - m1 : 0) + ((m2 != m1 && m2 < LT) ? LT - m2 : 0);
        m = n > m ? n : m;
    }
    return stage_stride(2 * m);
}
// the swap: each paired wave writes the sums of its partner's two pairs (rows of
// pairs G^1 and NPAIR-1-(G^1)) to its LDS slots and adds the partner's slots (the
// other source's sums of its own pairs) to its rows; 32 leaves at a time.  Every wave
// of the workgroup passes the same four barriers.
template <int L, int G>
__device__ __forceinline__ void sweep_block_swap(double2* stage, double (*ar)[L + 1], double (*ai)[L + 1]) {
    constexpr int LT = sweep_top_trunc<L>();
    constexpr int XS = swap_stride<L>();
    constexpr int P0 = G & ~1, Q0 = (L + 1) / 2 - 2 - P0;
    constexpr int p = G & 1;
    static_assert(((G | 1) + 1) * XS <= stage_stride(nhalf(L)), "the swap's slots exceed the staging buffer");
    // rows: 0 pair P0, 1 pair P0+1, 2 pair Q0, 3 pair Q0+1; wave P0 holds pairs P0 and
    // Q0+1, wave P0+1 pairs P0+1 and Q0
    constexpr int OWN0 = p ? 1 : 0, OWN1 = p ? 2 : 3, OTH0 = p ? 0 : 1, OTH1 = p ? 3 : 2;
    constexpr int OWN0P = p ? P0 + 1 : P0, OWN1P = p ? Q0 : Q0 + 1;
    constexpr int OTH0P = p ? P0 : P0 + 1, OTH1P = p ? Q0 + 1 : Q0;
    constexpr int N_OTH0 = pair_entries<L, OTH0P, LT>(), N_OWN0 = pair_entries<L, OWN0P, LT>();
    const int lane = threadIdx.x;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
        const bool mine = (lane >> 5) == half;
        if (mine) {
            double2* slot = stage + (G * 32 + (lane & 31)) * XS;
            swap_pair<L, OTH0P, LT, true>(slot, ar[OTH0], ai[OTH0]);
            swap_pair<L, OTH1P, LT, true>(slot + N_OTH0, ar[OTH1], ai[OTH1]);
        }
        __syncthreads();
        if (mine) {
            double2* slot = stage + ((G ^ 1) * 32 + (lane & 31)) * XS;
            swap_pair<L, OWN0P, LT, false>(slot, ar[OWN0], ai[OWN0]);
            swap_pair<L, OWN1P, LT, false>(slot + N_OWN0, ar[OWN1], ai[OWN1]);
        }
        __syncthreads();
    }
}
// truncated images in every bucket below L, from the block's lists
template <int L>
__device__ __forceinline__ int sweep_trunc_count(const int* lists) {
    int t = 0;
#pragma unroll
    for (int i = 0; i < SWEEP_NBUCKET; i++)
        if (sweep_bucket(i, L) < L) t += lists[SWEEP_NBUCKET * SWEEP_LIST + i];
    return __builtin_amdgcn_readfirstlane(t);
}

// the images of bucket I (truncation order lc in (bucket I-1, bucket I]) of both sources,
// in stencil order, for the lane's leaf
template <int L, int G, bool PER, int I>
__device__ __forceinline__ void sweep_block_images(const SweepArgs& a, const double* dx0, const double* dy0,
                                                   const double* dz0, const double* sq, const int* dcx,
                                                   const int* dcy, const int* dcz, const int* lists,
                                                   double (*ar)[L + 1], double (*ai)[L + 1]) {
    constexpr int LC = sweep_bucket(I, L);
    constexpr int LPREV = I > 0 ? sweep_bucket(I - 1, L) : 0;
    if constexpr (I < SWEEP_NBUCKET && (I == 0 || LPREV < L)) {
        if constexpr (G < LC) {
            constexpr int R = PER ? 1 : 0;
            const int n = a.g.ncps;
            const double E = a.g.E;
            const int cnt = __builtin_amdgcn_readfirstlane(lists[SWEEP_NBUCKET * SWEEP_LIST + I]);
#pragma unroll 1
            for (int i = 0; i < cnt; i++) {
                const int e = __builtin_amdgcn_readfirstlane(lists[I * SWEEP_LIST + i]);
                const int k = e & 255, lc = e >> 8;
                if (lc <= G) continue;   // (the image does not reach this wave's columns, its lowest is G)
                const int ox = k % (2 * R + 1) - R, oy = (k / (2 * R + 1)) % (2 * R + 1) - R,
                          oz = k / ((2 * R + 1) * (2 * R + 1)) - R;
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const double dx = fma((double)ox, E, dx0[s]), dy = fma((double)oy, E, dy0[s]),
                                 dz = fma((double)oz, E, dz0[s]);
                    if constexpr (LC < L) {   // truncated: no leaf of the block is a near neighbour
                        block_image<L, G, LC, false>(dx, dy, dz, sq[s], ar, ai);
                    } else {
                        const bool nr =
                            near1(dcx[s] + ox * n) && near1(dcy[s] + oy * n) && near1(dcz[s] + oz * n);
                        block_image<L, G, LC, true>(dx, dy, dz, nr ? 0.0 : sq[s], ar, ai);
                    }
                }
            }
        }
        sweep_block_images<L, G, PER, I + 1>(a, dx0, dy0, dz0, sq, dcx, dcy, dcz, lists, ar, ai);
    }
}

// group G (column pairs G*NP ..) of the move's 2 x nimg source images for the leaf
// (tx, ty, tz) of this lane; lists: the block's images by truncation bucket.  The
// images go bucket by bucket (each bucket's loop a compile-time truncation: unrolled,
// branch-free columns), in stencil order within a bucket.
template <int L, int G, int NP, int NI, bool PER>
__device__ __forceinline__ void sweep_block_group(const SweepArgs& a, int tx, int ty, int tz, const int* lists,
                                                  double2* stage, double (*ar)[L + 1], double (*ai)[L + 1]) {
    static_assert(NI == 1 && NP == 2, "the blocked sweep: one source image per pass, two column pairs per lane");
    const int n = a.g.ncps;
    const double cx = a.g.c0 + tx * a.g.w;
    const double cy = a.g.c0 + ty * a.g.w;
    const double cz = a.g.c0 + tz * a.g.w;
    const MoveRec* rc = a.rec;
    double dx0[2], dy0[2], dz0[2], sq[2];
    int dcx[2], dcy[2], dcz[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double* p = s == 0 ? rc->old_pos : rc->new_pos;
        const int sc = s == 0 ? rc->old_cell : rc->new_cell;
        sq[s] = s == 0 ? -1.0 * rc->q : rc->q;
        dx0[s] = p[0] - cx; dy0[s] = p[1] - cy; dz0[s] = p[2] - cz;
        dcx[s] = sc % n - tx; dcy[s] = (sc / n) % n - ty; dcz[s] = sc / (n * n) - tz;
    }
    if constexpr (block_paired<L, G>()) {   // (every wave of the workgroup, or none)
        // truncated buckets: source G & 1, four pairs, rows P0, P0+1, Q0, Q0+1; then the
        // swap; then this wave's own pairs (rows 0, 1) take the full bucket's images
        constexpr int p = G & 1;
        constexpr int W = L + 1;
        double tr[4][W], ti[4][W];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int k = 0; k < W; k++) { tr[j][k] = 0.0; ti[j][k] = 0.0; }
        sweep_block_trunc<L, G, PER, 0>(a, dx0[p], dy0[p], dz0[p], sq[p], lists, tr, ti);
        if (sweep_trunc_count<L>(lists) > 0) sweep_block_swap<L, G>(stage, tr, ti);
        constexpr int OWN0 = p ? 1 : 0, OWN1 = p ? 2 : 3;
#pragma unroll
        for (int k = 0; k < W; k++) {
            ar[0][k] = tr[OWN0][k]; ai[0][k] = ti[OWN0][k];
            ar[1][k] = tr[OWN1][k]; ai[1][k] = ti[OWN1][k];
        }
        sweep_block_images<L, G, PER, sweep_full_bucket<L>()>(a, dx0, dy0, dz0, sq, dcx, dcy, dcz, lists, ar, ai);
    } else {
        sweep_block_images<L, G, PER, 0>(a, dx0, dy0, dz0, sq, dcx, dcy, dcz, lists, ar, ai);
    }
}
// the scaled sums of wave G's columns for the lane's leaf, staged through LDS with the
// other waves' (two halves of 32 leaves), added to the block's expansions with
// coalesced loads and stores.  Only each leaf's first nh entries (the orders below the
// largest order any image of the block evaluates) are read and written: the others
// receive no term.
template <int L, int G>
__device__ __forceinline__ void sweep_block_update(const SweepArgs& a, double (*ar)[L + 1], double (*ai)[L + 1],
                                                   double2* stage, int cell_base, int nh) {
    constexpr int NH = nhalf(L);
    constexpr int ST = stage_stride(NH);
    const int lane = threadIdx.x;
    const int n = a.g.ncps;
    using PP = BlockPairs<L, G>;
    sweep_group_stage<L, PP::A, 1, true>(nullptr, ar, ai);
    if constexpr (PP::B > PP::A) sweep_group_stage<L, PP::B, 1, true>(nullptr, ar + 1, ai + 1);
    const int tid = threadIdx.x + 64 * threadIdx.y;
    const int nthr = 64 * blockDim.y;
#pragma unroll 1
    for (int half = 0; half < 2; half++) {
        __syncthreads();
        if ((lane >> 5) == half) {
            sweep_group_stage<L, PP::A, 1, false>(stage + (lane & 31) * ST, ar, ai);
            if constexpr (PP::B > PP::A) sweep_group_stage<L, PP::B, 1, false>(stage + (lane & 31) * ST, ar + 1, ai + 1);
        }
        __syncthreads();
        // the half's 32 leaves: 8 runs (y, z) of 4 consecutive leaves along x
        constexpr int U = 8;
        const int ne = 32 * nh;
#pragma unroll 1
        for (int e0 = tid; e0 < ne; e0 += nthr * U) {
            double2 v[U];
            size_t gi[U];
            int si[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * nthr;
                const int j = e / nh, h = e - j * nh;             // staged leaf, coefficient
                const int jy = (j >> 2) & 3, jz = 2 * half + (j >> 4);
                gi[u] = (size_t)(cell_base + (j & 3) + n * (jy + n * jz)) * NH + h;
                si[u] = j * ST + h;
                if (e < ne) v[u] = a.lexp[gi[u]];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (e0 + u * nthr < ne) {
                    const double2 sv = stage[si[u]];
                    v[u].x += sv.x;
                    v[u].y += sv.y;
                    a.lexp[gi[u]] = v[u];
                }
            }
        }
    }
}

// one wave's share of the blocked sweep: group G's columns of its lane's leaf, then the
// staged, coalesced update of the block's expansions together with the other waves
template <int L, int G, int NP, int NI, bool PER>
__device__ __forceinline__ void sweep_block_wave(const SweepArgs& a, int tx, int ty, int tz, const int* lists,
                                                 double2* stage, int cell_base, int nh) {
    constexpr int W = L + 1;
    double ar[NP][W], ai[NP][W];
#pragma unroll
    for (int j = 0; j < NP; j++)
#pragma unroll
        for (int k = 0; k < W; k++) { ar[j][k] = 0.0; ai[j][k] = 0.0; }
    sweep_block_group<L, G, NP, NI, PER>(a, tx, ty, tz, lists, stage, ar, ai);
    sweep_block_update<L, G>(a, ar, ai, stage, cell_base, nh);
}
template <int L, int NP, int NI, bool PER, int G0>
__device__ __forceinline__ void sweep_block_dispatch(const SweepArgs& a, int group, int tx, int ty, int tz,
                                                     const int* lists, double2* stage, int cell_base, int nh) {
    constexpr int NG = ((L + 1) / 2 + NP - 1) / NP;
    if constexpr (G0 < NG) {
        if (group == G0) sweep_block_wave<L, G0, NP, NI, PER>(a, tx, ty, tz, lists, stage, cell_base, nh);
        else sweep_block_dispatch<L, NP, NI, PER, G0 + 1>(a, group, tx, ty, tz, lists, stage, cell_base, nh);
    }
}

// truncation order of source image k (both sources) against the 4x4x4 block (bx, by, bz):
// the smallest order whose tail bound is below SWEEP_TAU for every leaf of the block
template <int L, bool PER>
__device__ __forceinline__ int sweep_block_lc(const SweepArgs& a, int bx, int by, int bz, int k) {
    const int R = PER ? 1 : 0;
    const int ox = k % (2 * R + 1) - R, oy = (k / (2 * R + 1)) % (2 * R + 1) - R,
              oz = k / ((2 * R + 1) * (2 * R + 1)) - R;
    const double w = a.g.w;
    const double rho = 0.8660254037844386 * w;   // half diagonal of a leaf
    const double hb = 1.5 * w;                    // the 4 leaf centres along an axis span 3w
    const double mx = a.g.c0 + (4 * bx + 1.5) * w, my = a.g.c0 + (4 * by + 1.5) * w,
                 mz = a.g.c0 + (4 * bz + 1.5) * w;
    int lc = 0;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double* p = s == 0 ? a.rec->old_pos : a.rec->new_pos;
        const double ex = fmax(fabs(p[0] + ox * a.g.E - mx) - hb, 0.0);
        const double ey = fmax(fabs(p[1] + oy * a.g.E - my) - hb, 0.0);
        const double ez = fmax(fabs(p[2] + oz * a.g.E - mz) - hb, 0.0);
        lc = max(lc, sweep_trunc_order(sqrt(fma(ex, ex, fma(ey, ey, ez * ez))), rho, L));
    }
    return lc;
}

// grid: one workgroup per 4x4x4 block of leaves (ncps a multiple of 4), 64 x NG threads
template <int L, int NP, int NI, bool PER>
__device__ __forceinline__ void sweep_block_body(const SweepArgs& a, int b) {
    constexpr int NH = nhalf(L);
    constexpr int ST = stage_stride(NH);
    constexpr int NIMG = PER ? 27 : 1;
    __shared__ double2 stage[32 * ST];
    __shared__ int s_list[SWEEP_NBUCKET * SWEEP_LIST + SWEEP_NBUCKET];
    __shared__ int s_nh;
    const int lane = threadIdx.x;
    const int group = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int n = a.g.ncps, nb = n >> 2;
    const int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
    const int tx = 4 * bx + (lane & 3), ty = 4 * by + ((lane >> 2) & 3), tz = 4 * bz + (lane >> 4);
    if (!a.rec->valid) return;   // (uniform over the grid: nothing to add)
    if (group == 0) {
        // lane k < NIMG: image k's truncation order; its bucket's list gets k at its rank
        // among the bucket's images (stencil order)
        int lc = lane < NIMG ? sweep_block_lc<L, PER>(a, bx, by, bz, lane) : 0;
        const int bi = lc > 0 ? sweep_bucket_index(lc, L) : -1;
        const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
        for (int I = 0; I < SWEEP_NBUCKET; I++) {
            const uint64_t m = __ballot(bi == I);
            if (bi == I) s_list[I * SWEEP_LIST + __popcll(m & below)] = lane | (lc << 8);
            if (lane == 0) s_list[SWEEP_NBUCKET * SWEEP_LIST + I] = __popcll(m);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) lc = max(lc, __shfl_xor(lc, o, 64));
        // expansion entries that receive terms: the orders below the largest evaluated
        if (lane == 0) s_nh = nhalf(sweep_bucket_of(lc, L));
    }
    __syncthreads();
    const int nh = __builtin_amdgcn_readfirstlane(s_nh);
    if (nh == 0) return;   // (uniform over the workgroup)
    const int cell_base = 4 * bx + n * (4 * by + n * 4 * bz);
    sweep_block_dispatch<L, NP, NI, PER, 0>(a, group, tx, ty, tz, s_list, stage, cell_base, nh);
}

template <int L, int NP, int NI, bool PER>
__global__ void __launch_bounds__(512, 2) k_sweep_block(SweepArgs a) {
    sweep_block_body<L, NP, NI, PER>(a, (int)blockIdx.x);
}
// batches: one flat grid over the (system, block) pairs of a group, tab[blockIdx.x] =
// (system index in the argument table, block of that system), so no workgroup is
// launched for blocks a smaller system does not have
template <int L, int NP, int NI, bool PER>
__global__ void __launch_bounds__(512, 2) kb_sweep_block(const SweepArgs* __restrict__ args,
                                                         const int2* __restrict__ tab) {
    const int2 e = tab[blockIdx.x];
    sweep_block_body<L, NP, NI, PER>(args[e.x], e.y);
}


template <int L, int NP, bool PER>
__global__ void __launch_bounds__(512, ((NP == 1 && L <= 16) ? 4 : 2)) k_sweep_fused(SweepArgs a) {
    sweep_fused_body<L, NP, PER>(a);
}
template <int L, int NP, bool PER>
__global__ void __launch_bounds__(512, ((NP == 1 && L <= 16) ? 4 : 2)) kb_sweep_fused(const SweepArgs* __restrict__ args) {
    const SweepArgs& a = args[blockIdx.z];
    if ((int)blockIdx.x * 64 >= a.ncells) return;   // grid sized for the largest system of the batch
    sweep_fused_body<L, NP, PER>(a);
}


// ============================================================================
// far field: phi = Re sum L_lm R_l^m(x - centre), one lane per evaluation point
// ============================================================================
struct EvalArgs {
    Geom g;
    int n;                       // evaluation points
    const double4* pos4;         // (x, y, z, q)
    const int32_t* cell;
    const int32_t* excl;         // id excluded from the near field (-1: none)
    double* out;
    const double2* lexp;
    const int* occ;
    const double4* slot_pos;
    const int32_t* slot_id;
    int occmax;
    const double4* ppos4;        // particle positions (x, y, z, q) by id
    int np;                      // particles
};

// columns M1 and M2 (M2 == M1: one column) of  Re sum_l L_lm R_l^m(d)  with weight
// 2 for m > 0 (the m < 0 terms are the complex conjugates)
template <int L, int M1, int M2>
__device__ __forceinline__ double far_columns(double dx, double dy, double dz, double r2, const double2* le) {
    CSum phi;
    const Cplx w = {dx, dy};
    const Cplx C1 = cpow<M1>(w);
#pragma unroll
    for (int col = 0; col < 2; col++) {
        if (col == 1 && M2 == M1) break;
        const int m = col == 0 ? M1 : M2;
        const Cplx C = col == 0 ? C1 : cmul(C1, cpow<M2 - M1>(w));
        double b2 = 0.0, b1 = kRec.s[m];
        double2 v = le[hidx(m, m)];
        CSum sr, si;
        sr.add_prod(b1, v.x);
        si.add_prod(b1, v.y);
#pragma unroll
        for (int l = m + 1; l < L; l++) {
            const double bb = fma(kRec.a[l][m] * dz, b1, -(kRec.b[l][m] * r2) * b2);
            b2 = b1;
            b1 = bb;
            v = le[hidx(l, m)];
            sr.add_prod(bb, v.x);
            si.add_prod(bb, v.y);
        }
        // Re(C (sr + i si)) with weight 2 for m > 0 (exact scaling)
        const double wgt = m == 0 ? 1.0 : 2.0;
        phi.add_prod(wgt * C.re, sr.s);
        phi.add_prod(-wgt * C.im, si.s);
        phi.add(wgt * fma(C.re, sr.c, -C.im * si.c));
    }
    return phi.value();
}

template <int L, int P0>
__device__ __forceinline__ double far_dispatch(int pair, double dx, double dy, double dz, double r2,
                                               const double2* le) {
    if constexpr (P0 < (L + 1) / 2) {
        if (pair == P0) return far_columns<L, P0, L - 1 - P0>(dx, dy, dz, r2, le);
        return far_dispatch<L, P0 + 1>(pair, dx, dy, dz, r2, le);
    } else {
        return 0.0;
    }
}

// one lane per evaluation point; the column pairs (m, L-1-m) of 64 points are the
// waves of one workgroup (pair = blockIdx.y * blockDim.y + threadIdx.y), so the
// waves reading one leaf expansion run together and share its cache lines.  The
// workgroup's pair terms are summed in pair order with a compensated sum (through LDS)
// and written as its (sum, correction): out[(2 blockIdx.y + {0, 1}) n + i], so a point
// has far_parts(L) partials (2 up to L = 16, 4 above).
constexpr int FAR_PY_MAX = 8;   // pair waves per workgroup
__host__ __device__ constexpr int far_pairs_per_group(int L) {
    return (L + 1) / 2 <= FAR_PY_MAX ? (L + 1) / 2 : ((L + 1) / 2 + 1) / 2;
}
__host__ __device__ constexpr int far_parts(int L) {
    return 2 * (((L + 1) / 2 + far_pairs_per_group(L) - 1) / far_pairs_per_group(L));
}
template <int L>
__device__ __forceinline__ void far_body(const EvalArgs& a) {
    __shared__ double s_part[FAR_PY_MAX][64];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int pair = __builtin_amdgcn_readfirstlane((int)(blockIdx.y * blockDim.y + threadIdx.y));
    double v = 0.0;
    if (i < a.n && pair < (L + 1) / 2) {
        const double4 p = a.pos4[i];
        const int c = a.cell[i];
        const int n = a.g.ncps;
        const int tx = c % n, ty = (c / n) % n, tz = c / (n * n);
        const double dx = p.x - (a.g.c0 + tx * a.g.w);
        const double dy = p.y - (a.g.c0 + ty * a.g.w);
        const double dz = p.z - (a.g.c0 + tz * a.g.w);
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        const double2* le = a.lexp + (size_t)c * nhalf(L);
        v = p.w * far_dispatch<L, 0>(pair, dx, dy, dz, r2, le);
    }
    s_part[threadIdx.y][threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.y == 0 && i < a.n) {
        CSum acc;
        for (int w = 0; w < (int)blockDim.y; w++) acc.add(s_part[w][threadIdx.x]);
        a.out[(size_t)(2 * blockIdx.y) * a.n + i] = acc.s;
        a.out[(size_t)(2 * blockIdx.y + 1) * a.n + i] = acc.c;
    }
}
template <int L>
__global__ void __launch_bounds__(512) k_far(EvalArgs a) { far_body<L>(a); }
template <int L>
__global__ void __launch_bounds__(512) kb_far(const EvalArgs* __restrict__ args) { far_body<L>(args[blockIdx.z]); }


// ============================================================================
// near field: q_t * sum over the 27 neighbour leaves of q_j / |x_t - x_j|
// G lanes per evaluation point, neighbour slots strided over the G lanes,
// shuffle reduction inside the lane group.
// ============================================================================
// lane `lane` of the G lanes of evaluation entry i: its share of the slots of the 27
// leaves around the entry's leaf
template <int G, bool MIXED>
__device__ __forceinline__ CSum near_point(const EvalArgs& a, int i, int lane) {
    // The excluded charge (the moving one, for a moving charge's old position) is the
    // one slot at distance exactly 0 from the point: its slot holds exactly the point's
    // position, and no other charge can sit at that same point (their pair energy would
    // be infinite).  For such points r2 gets 1e-30 added (below the rounding of any
    // other pair's r2, and a normal FP32 number for the mixed 1/sqrt seed), so that pair
    // stays finite, and its charge is taken as 0: every term is then added without a
    // branch, the excluded one as an exact 0 (the compensated sum is unchanged by it).
    // No slot id is loaded: a 4-byte id load touches as many cache lines as the 32-byte
    // record load.
    CSum acc;
    {
        const double4 p = a.pos4[i];
        const int c = a.cell[i];
        const int ex = a.excl[i];
        const bool has_ex = ex >= 0 && ex < a.np;
        const double tiny = has_ex ? 1e-30 : 0.0;
        const int n = a.g.ncps;
        const int tx = c % n, ty = (c / n) % n, tz = c / (n * n);
        const bool periodic = a.g.bc != KMC_BC_FREE_SPACE;
        // per axis, the leaf index (wrapped, or -1 outside a free-space box) and the
        // image shift of offsets -1, 0, +1, worked out once per point; the 27 leaves are
        // then walked in stencil order (x fastest, then y, then z) with no per-leaf
        // wrapping
        int xi[3], yi[3], zi[3];
        double xs[3], ys[3], zs[3];
        auto axis = [&](int t, int* id, double* sh) {
#pragma unroll
            for (int o = 0; o < 3; o++) {
                int v = t + o - 1;
                double h = 0.0;
                if (periodic) {
                    if (v < 0) { v += n; h = -a.g.E; } else if (v >= n) { v -= n; h = a.g.E; }
                } else if (v < 0 || v >= n) {
                    v = -1;
                }
                id[o] = v;
                sh[o] = h;
            }
        };
        axis(tx, xi, xs);
        axis(ty, yi, ys);
        axis(tz, zi, zs);
#pragma unroll
        for (int kz = 0; kz < 3; kz++) {
#pragma unroll
            for (int ky = 0; ky < 3; ky++) {
                const int row = n * (yi[ky] + n * zi[kz]);
                const double py = p.y - ys[ky], pz = p.z - zs[kz];
#pragma unroll
                for (int kx = 0; kx < 3; kx++) {
                    if (xi[kx] < 0 || yi[ky] < 0 || zi[kz] < 0) continue;   // free space: outside the box
                    const int nc = xi[kx] + row;
                    const int occ = min(a.occ[nc], a.occmax);
                    const double px = p.x - xs[kx];
                    const size_t base = (size_t)nc * a.occmax;
                    for (int s = lane; s < occ; s += G) {
                        const double4 q = a.slot_pos[base + s];
                        const double dx = px - q.x, dy = py - q.y, dz = pz - q.z;
                        const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, tiny)));
                        const double qs = (has_ex && r2 == tiny) ? 0.0 : q.w;
                        acc.add(qs * rsqrt_sel<MIXED>(r2));
                    }
                }
            }
        }
    }
    return acc;
}
template <int G, bool MIXED>
__device__ __forceinline__ void near_body(const EvalArgs& a) {
    const int gid = (blockIdx.x * blockDim.x + threadIdx.x);
    const int i = gid / G;
    const int lane = gid % G;
    const bool active = i < a.n;
    CSum acc;
    if (active) acc = near_point<G, MIXED>(a, i, lane);
    if constexpr (G > 1) acc = csum_group<G>(acc);
    if (active && lane == 0) a.out[i] = acc.value() * a.pos4[i].w;
}

// The per-point near field over the entries in leaf-tile order (perm: the counting sort
// of the entries by 4x4x4 tile of leaves).  Entry perm[i] is summed exactly as near_body
// sums entry perm[i], so the results equal k_near<G>'s bit for bit; consecutive entries
// share most of their 27 leaves.  Workgroups go to the XCDs round-robin (b, b + 8, ... on
// one XCD), so block b takes logical block lb, which gives each XCD one contiguous range
// of the sorted entries and keeps their shared leaves in that XCD's L2.
template <int G, bool MIXED>
__device__ __forceinline__ void near_sorted_body(const EvalArgs& a, const int* __restrict__ perm) {
    const int nblk = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nblk >> 3, r = nblk & 7, x = b & 7, slot = b >> 3;
    const int lb = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
    const int gid = lb * (int)blockDim.x + (int)threadIdx.x;
    const int i = gid / G;
    const int lane = gid % G;
    const bool active = i < a.n;
    const int e = active ? perm[i] : 0;
    CSum acc;
    if (active) acc = near_point<G, MIXED>(a, e, lane);
    if constexpr (G > 1) acc = csum_group<G>(acc);
    if (active && lane == 0) a.out[e] = acc.value() * a.pos4[e].w;
}

// near field, one lane per neighbour leaf: 32 lanes per evaluation point, lanes
// 0..26 each walk one leaf of the stencil, shuffle reduction over the 32 lanes.
template <bool MIXED>
__device__ __forceinline__ void near_cells_body(const EvalArgs& a) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gid >> 5;
    const int k = gid & 31;
    CSum acc;
    if (i < a.n && k < 27) {
        const double4 p = a.pos4[i];
        const int c = a.cell[i];
        const int ex = a.excl[i];
        const int n = a.g.ncps;
        const int tx = c % n, ty = (c / n) % n, tz = c / (n * n);
        int ox, oy, oz;
        stencil(k, ox, oy, oz);
        int nx = tx + ox, ny = ty + oy, nz = tz + oz;
        double shx = 0.0, shy = 0.0, shz = 0.0;
        bool ok = true;
        if (a.g.bc != KMC_BC_FREE_SPACE) {
            if (nx < 0) { nx += n; shx = -a.g.E; } else if (nx >= n) { nx -= n; shx = a.g.E; }
            if (ny < 0) { ny += n; shy = -a.g.E; } else if (ny >= n) { ny -= n; shy = a.g.E; }
            if (nz < 0) { nz += n; shz = -a.g.E; } else if (nz >= n) { nz -= n; shz = a.g.E; }
        } else if (nx < 0 || nx >= n || ny < 0 || ny >= n || nz < 0 || nz >= n) {
            ok = false;
        }
        if (ok) {
            const int nc = nx + n * (ny + n * nz);
            const int occ = min(a.occ[nc], a.occmax);
            const double px = p.x - shx, py = p.y - shy, pz = p.z - shz;
            const size_t base = (size_t)nc * a.occmax;
            // slots in groups of 4 with all loads issued up front (indices clamped to
            // the cell's slot range; slots past the occupancy contribute 0)
            for (int s0 = 0; s0 < occ; s0 += 4) {
                double4 q[4];
#pragma unroll
                for (int j = 0; j < 4; j++) q[j] = a.slot_pos[base + min(s0 + j, a.occmax - 1)];
                int id[4] = {-1, -1, -1, -1};
                if (ex >= 0) {
#pragma unroll
                    for (int j = 0; j < 4; j++) id[j] = a.slot_id[base + min(s0 + j, a.occmax - 1)];
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const double dx = px - q[j].x, dy = py - q[j].y, dz = pz - q[j].z;
                    double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    const bool use = (s0 + j < occ) && (ex < 0 || id[j] != ex);
                    r2 = use ? r2 : 1.0;
                    const double contrib = q[j].w * rsqrt_sel<MIXED>(r2);
                    acc.add(use ? contrib : 0.0);
                }
            }
        }
    }
    acc = csum_group<32>(acc);
    if (i < a.n && k == 0) a.out[i] = acc.value() * a.pos4[i].w;
}

// near field of the propose paths, grouped by moving charge: one wave per moving
// charge p takes its old position (entry T + p, its own id excluded) and its proposals
// (entries es[p] .. es[p+1]-1) together, up to 16 points per pass.  The leaves within
// one of the points' leaves form a box around the old position's leaf (at most 5 per
// side); the wave lists the box's slots leaf by leaf (prefix sums in LDS) and each lane
// reads its share of them once and adds each to every point whose 27-leaf stencil holds
// its leaf.  Per pair the image shift and the arithmetic are those of the per-point
// kernels (only the order of the sums differs).  A pass whose points span more than
// 3 leaves along an axis is evaluated point by point over each point's own 27 leaves.
__device__ __forceinline__ double readlane_d(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ int floordiv(int a, int n) { return a >= 0 ? a / n : -((-a + n - 1) / n); }

constexpr int NEAR_GROUP_POINTS = 16;
constexpr int NEAR_BOX = 5;



// ============================================================================
// proposal setup (tuple form) and validation
// ============================================================================
struct SetupArgs {
    Geom g;
    double max_move;
    int P, T, n;
    const int64_t* pid;
    const int64_t* es;
    const double* new_pos;
    const double* new_q;
    const double4* ppos4;
    const int32_t* pcell;
    double4* epos4;
    int32_t* ecell;
    int32_t* eexcl;
    int32_t* owner;
    int* err;
    int64_t* es_out;     // copy of es[0..P] (read by the grouped near-field kernel)
    const int64_t* mask; // proposals with mask <= 0 are not validated (they get rate 0), or null
};

// owner p of proposal t: es[p] <= t < es[p+1]; always returns a value in [0, P)
__device__ __forceinline__ int owner_of(int t, const int64_t* es, int P) {
    int lo = 0, hi = P - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (es[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One thread per proposal t (< T) and per moving charge p (< P).
// kmc_mpi_decomp.py:312-345 (domain and max_move checks, minimum image for '27'/PBC)
__device__ __forceinline__ void setup_body(const SetupArgs& a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    if (x < a.P) {
        int i = (int)a.pid[x];
        if (i < 0 || i >= a.n) { bad++; i = 0; }
        a.epos4[a.T + x] = a.ppos4[i];
        a.ecell[a.T + x] = a.pcell[i];
        a.eexcl[a.T + x] = i;
        if (a.es[x] > a.es[x + 1]) bad++;
        if (x == 0 && a.es[0] != 0) bad++;
        if (x == a.P - 1 && a.es[a.P] != a.T) bad++;
        a.es_out[x] = a.es[x];
        if (x == a.P - 1) a.es_out[a.P] = a.es[a.P];
    }
    if (x < a.T) {
        const int p = owner_of(x, a.es, a.P);
        a.owner[x] = p;
        int i = (int)a.pid[p];
        if (i < 0 || i >= a.n) i = 0;
        const double4 o = a.ppos4[i];
        const int n = a.g.ncps;
        const double px = a.new_pos[3 * x], py = a.new_pos[3 * x + 1], pz = a.new_pos[3 * x + 2];
        const double nq = a.new_q ? a.new_q[x] : o.w;
        a.epos4[x] = make_double4(px, py, pz, nq);
        const int cx = bin_axis(px, a.g.E, a.g.w, n), cy = bin_axis(py, a.g.E, a.g.w, n), cz = bin_axis(pz, a.g.E, a.g.w, n);
        a.ecell[x] = cx + n * (cy + n * cz);
        a.eexcl[x] = -1;
        const double np3[3] = {px, py, pz};
        const double op3[3] = {o.x, o.y, o.z};
        const int nd = (a.mask && a.mask[x] <= 0) ? 0 : 3;
        for (int d = 0; d < nd; d++) {
            if (!(np3[d] >= -0.5 * a.g.E && np3[d] <= 0.5 * a.g.E)) bad++;
            double dd = np3[d] - op3[d];
            if (a.g.bc != KMC_BC_FREE_SPACE) {
                if (dd < -0.5 * a.g.E) dd += a.g.E;
                if (dd > 0.5 * a.g.E) dd -= a.g.E;
            }
            if (fabs(dd) > a.max_move) bad++;
        }
    }
    if (bad) atomicAdd(a.err, bad);
}


// ============================================================================
// self term + long-range difference + combine (diff_lib)
// ============================================================================
struct CombineArgs {
    Geom g;
    int T;
    double eu;
    double self27;               // -sum_{o != 0} 1/|o E|
    const double4* epos4;
    const int32_t* owner;
    const double* UD;
    const double* UI;            // far-field partials [nfar][T+P] (nfar = far_parts(L))
    int nfar;
    const double* lpart;         // long range per proposal partials [12][T], or null
    const double* cpart;         // long range per moving charge partials [ncpart][P]
    int ncpart;                  // 12 x the long range's row chunks
    int P;
    const int64_t* rate_loc;     // output index, or null (identity)
    double* out;
    double kT;
    double* rates;               // rates per proposal (rate_rule), or null
    double* bsum;                // their sums over blocks of 64 proposals
    const int64_t* mask;         // proposals with mask <= 0 get rate 0, or null
    int rate_rule;               // 0: min(1, exp(-dE/kT)) (Metropolis), 1: exp(-dE/kT)
};

// Partial sums v[j * stride + i], j < n, loaded into registers before they are added:
// each load of the chunk is issued before the first add, so a lane waits on memory once
// per chunk rather than once per partial (n is a kernel argument: a plain loop would
// wait on each load in turn).
template <int C>
__device__ __forceinline__ void load_parts(double (&buf)[C], const double* v, int j0, int n, size_t stride,
                                           size_t i) {
#pragma unroll
    for (int j = 0; j < C; j++) buf[j] = j0 + j < n ? v[(size_t)(j0 + j) * stride + i] : 0.0;
}
template <int C>
__device__ __forceinline__ void add_loaded(CSum& acc, double sign, const double (&buf)[C], int j0, int n) {
#pragma unroll
    for (int j = 0; j < C; j++)
        if (j0 + j < n) acc.add(sign * buf[j]);
}

// Metropolis (rule 0) or exponential (rule 1) rate of an energy change.  An exponential
// rate beyond 1e300 (dE below -690 kT) is held at 1e300: it still dominates the
// selection, and a sum of up to ~1e8 such rates stays finite.
__device__ __forceinline__ double rate_of(double de, double kT, int rule) {
    if (!(de == de) || isinf(de)) return 0.0;
    const double w = exp(-de / kT);
    if (rule == 1) return w < 1e300 ? w : 1e300;
    return w < 1.0 ? w : 1.0;
}

// energy difference of proposal t (writes it at its output location):
//   dE = eu [ (U1D + U1I) - (U0D + U0I) - USI ],  USI = q_o^2 e27 - (long-range parts)
// (diff_lib, coulomb_kmc/kmc_fmm.py:583-663), every term added into one compensated
// sum, so the old - new cancellation costs no more than one rounding of dE itself.
template <bool MIXED>
__device__ __forceinline__ double combine_one(const CombineArgs& a, int t) {
    constexpr int NF = (KMAX_L + 1) / 2;      // far-field partials at most
    const int ow = a.owner[t];
    const int po = a.T + ow;
    const double4 xo = a.epos4[po];
    const double4 xn = a.epos4[t];
    // every partial loaded up front (first 12 per-charge long-range parts included); the
    // self term below runs while the loads are in flight
    const bool direct = a.g.bc != KMC_BC_FF_ONLY;
    const size_t ne = (size_t)a.T + a.P;
    double un = 0.0, uo = 0.0, fn[NF], fo[NF], lp[12], cp[12];
    if (direct) {
        un = a.UD[t];
        uo = a.UD[po];
        load_parts(fn, a.UI, 0, a.nfar, ne, t);
        load_parts(fo, a.UI, 0, a.nfar, ne, po);
    }
    if (a.lpart) {
        load_parts(lp, a.lpart, 0, 12, a.T, t);
        load_parts(cp, a.cpart, 0, a.ncpart, a.P, ow);
    }
    // self term (kmc_fmm_self_interaction.py:48-54, 163-182, 227-236)
    CSum e27;
    if (a.g.bc == KMC_BC_FREE_SPACE) {
        const double dx = xo.x - xn.x, dy = xo.y - xn.y, dz = xo.z - xn.z;
        e27.add(rsqrt_sel<MIXED>(fma(dx, dx, fma(dy, dy, dz * dz))));
    } else if (a.g.bc != KMC_BC_FF_ONLY) {
        e27.add(a.self27);
        for (int k = 0; k < 27; k++) {
            int ox, oy, oz;
            stencil(k, ox, oy, oz);
            const double dx = ox * a.g.E + xo.x - xn.x;
            const double dy = oy * a.g.E + xo.y - xn.y;
            const double dz = oz * a.g.E + xo.z - xn.z;
            e27.add(rsqrt_sel<MIXED>(fma(dx, dx, fma(dy, dy, dz * dz))));
        }
    }
    CSum tot;
    if (direct) {
        tot.add(un);
        add_loaded(tot, 1.0, fn, 0, a.nfar);
        tot.add(-uo);
        add_loaded(tot, -1.0, fo, 0, a.nfar);
    }
    const double q2 = xo.w * xo.w;
    tot.add_prod(-q2, e27.s);
    tot.add(-q2 * e27.c);
    if (a.lpart) {
        add_loaded(tot, 1.0, lp, 0, 12);
        add_loaded(tot, 1.0, cp, 0, a.ncpart);
        for (int j0 = 12; j0 < a.ncpart; j0 += 12) {   // row-split long range: 12 per chunk
            load_parts(cp, a.cpart, j0, a.ncpart, a.P, ow);
            add_loaded(tot, 1.0, cp, j0, a.ncpart);
        }
    }
    const double de = a.eu * tot.value();
    const int64_t o = a.rate_loc ? a.rate_loc[t] : (int64_t)t;
    a.out[o] = de;
    return de;
}

// one lane per proposal, blocks of 64; with rates: the Metropolis rate per proposal
// and the rate sum of each block of 64 (for the selection)
template <bool MIXED>
__device__ __forceinline__ void combine_body(const CombineArgs& a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double rate = 0.0;
    if (t < a.T) {
        const double de = combine_one<MIXED>(a, t);
        if (a.rates) {
            rate = (a.mask && a.mask[t] <= 0) ? 0.0 : rate_of(de, a.kT, a.rate_rule);
            a.rates[t] = rate;
        }
    }
    if (a.rates && (int)blockIdx.x * 64 < a.T) {   // (a batched grid can run past this system's T)
        const double bs = wave_sum(rate);
        if (threadIdx.x == 0) a.bsum[blockIdx.x] = bs;
    }
}


// ============================================================================
// long range
// ============================================================================
// packed multipole of charge q at x about the origin: M_lm = q conj(R_l^m(x)), m >= 0.
// Lane m (< L) writes column m into dst (stride 2 doubles per (l,m)).
template <int L>
__device__ __forceinline__ void multipole_column(int m, double x, double y, double z, double q, double* dst,
                                                 double sign, bool accumulate) {
    const double r2 = fma(x, x, fma(y, y, z * z));
    Cplx C = {1.0, 0.0};
    const Cplx w = {x, -y};
    for (int k = 0; k < m; k++) C = cmul(C, w);
    double b2 = 0.0, b1 = 0.0;
#pragma unroll
    for (int l = 0; l < L; l++) {
        if (l < m) continue;
        double b;
        if (l == m) b = kRec.s[m];
        else b = fma(kRec.a[l][m] * z, b1, -(kRec.b[l][m] * r2) * b2);
        b2 = b1;
        b1 = b;
        const double vr = sign * q * b * C.re, vi = sign * q * b * C.im;
        const int h = hidx(l, m);
        if (accumulate) {
            dst[2 * h] += vr;
            dst[2 * h + 1] += vi;
        } else {
            dst[2 * h] = vr;
            dst[2 * h + 1] = vi;
        }
    }
}

// ---------------------------------------------------------------------------
// Long range, blocked form.  Under the cubic lattice's symmetry G is exactly
// block diagonal over 12 blocks: (real | imaginary part) x (m mod 4 class:
// {0}, {2}, odd) x (parity of l) - only lattice sums S_p^q with p even and
// q = 0 mod 4 survive.  Entries of a block are ordered by m, then l.  The host
// gathers the dense blocks (kmc_set_lr_operator); the kernels below keep a
// block's multipole entries in registers and read the block coefficients as
// wave-uniform (scalar) operands.
//
//   dLR(old - new) = -( v_n . a_o + 0.5 v_n G v_n + c_o ),
//   a_o = g - G v_o,  c_o = 0.5 v_o G v_o - g . v_o,  g = G M,
// with v = q conj(R(x)) the packed multipole of the moved charge.
// ---------------------------------------------------------------------------
constexpr int LRB_MAX = 64;
struct LRBlocks {
    int n[6];
    int off[6];            // offset of the block pair (real block, then imaginary block) in gblk
    int l[6][LRB_MAX];
    int m[6][LRB_MAX];
};
__host__ __device__ constexpr int m_class(int m) { return (m % 4 == 0) ? 0 : ((m % 4 == 2) ? 1 : 2); }
__host__ __device__ constexpr LRBlocks make_lr_blocks(int L) {
    LRBlocks t{};
    int off = 0;
    for (int b = 0; b < 6; b++) {
        const int c = b / 2, par = b % 2;
        int k = 0;
        for (int m = 0; m < L; m++) {
            if (m_class(m) != c) continue;
            for (int l = m; l < L; l++)
                if ((l & 1) == par) {
                    t.l[b][k] = l;
                    t.m[b][k] = m;
                    k++;
                }
        }
        t.n[b] = k;
        t.off[b] = off;
        off += 2 * k * k;
    }
    return t;
}
template <int L>
struct LRTab {
    static constexpr LRBlocks T = make_lr_blocks(L);
};

// entries of block B of v = q conj(R(x)): vr = real parts, vi = imaginary parts, from
// one column recurrence b_l^m per m of the block's class (values bb * q Re/Im(C_m))
template <int L, int B>
__device__ __forceinline__ void lr_block_values2(double x, double y, double z, double q, double* vr, double* vi) {
    constexpr int C = B / 2, PAR = B % 2;
    const double r2 = fma(x, x, fma(y, y, z * z));
    const Cplx w = {x, -y};
    const Cplx w2 = cmul(w, w);
    const Cplx w4 = cmul(w2, w2);
    Cplx Cm = (C == 0) ? Cplx{1.0, 0.0} : ((C == 1) ? w2 : w);
    const Cplx step = (C == 2) ? w2 : w4;
    int k = 0;
#pragma unroll
    for (int m = 0; m < L; m++) {
        if (m_class(m) != C) continue;
        const double qr = q * Cm.re, qi = q * Cm.im;
        double b2 = 0.0, b1 = 0.0;
#pragma unroll
        for (int l = m; l < L; l++) {
            const double bb = (l == m) ? kRec.s[m] : fma(kRec.a[l][m] * z, b1, -(kRec.b[l][m] * r2) * b2);
            b2 = b1;
            b1 = bb;
            if ((l & 1) == PAR) {
                vr[k] = bb * qr;
                vi[k] = bb * qi;
                k++;
            }
        }
        Cm = cmul(Cm, step);
    }
}

__host__ __device__ constexpr int lr_nbmax(int L) {
    const LRBlocks t = make_lr_blocks(L);
    int m = 0;
    for (int b = 0; b < 6; b++) m = t.n[b] > m ? t.n[b] : m;
    return m + (m & 1);   // even: rows of a_o stay 16-byte aligned
}

struct LR2Args {
    int P, T;
    const double* gblk;
    const double* gvec;
    const double4* epos4;
    const int32_t* owner;
    double* aold;              // [12][NBMAX][P]  a_o per (block, component), entry-major: the lanes
                               // (moving charges / their proposals) read and write adjacent words
    double* cpart;             // [12][rsplit][P]   c_o per (block, component, row chunk)
    int rsplit;                // row chunks of each block in k_lr_part (1..LR_SPLIT_MAX)
    double* lpart;             // [12][T]   v.a_o + 0.5 v G v per (block, component)
};

// per moving charge p, both components of block B, rows [i0, i1) of chunk r of the
// block's a.rsplit row chunks: a_o = g - G v_o, and the chunk's share of c_o.  With few
// moving charges the chunks give the kernel enough lanes to fill the device.  The
// block's coefficients are wave-uniform: read straight from memory they become scalar
// loads (s_load into SGPRs), one per several FMAs.
constexpr int LR_SPLIT_MAX = 4;
template <int L, int B>
__device__ __forceinline__ void lr_block_part(const LR2Args& a, const double* __restrict__ gblk,
                                              const double* __restrict__ g, int r) {
    constexpr LRBlocks T = LRTab<L>::T;
    constexpr int NB = T.n[B];
    constexpr int NBMAX = lr_nbmax(L);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P || r >= a.rsplit) return;
    const int R = a.rsplit;
    if constexpr (NB > 0) {
        const int i0 = r * NB / R, i1 = (r + 1) * NB / R;
        const double4 xo = a.epos4[a.T + p];
        double vr[NB], vi[NB];
        lr_block_values2<L, B>(xo.x, xo.y, xo.z, xo.w, vr, vi);
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            const double* __restrict__ G = gblk + T.off[B] + comp * NB * NB;
            const double* v = comp == 0 ? vr : vi;
            const int Y = 2 * B + comp;
            double* ao = a.aold + (size_t)Y * NBMAX * a.P + p;   // entry i at ao[i * P]
            double c = 0.0;
#pragma unroll
            for (int i = 0; i < NB; i++) {
                if (i < i0 || i >= i1) continue;
                const int h = 2 * hidx(T.l[B][i], T.m[B][i]) + comp;
                double yy = 0.0;
#pragma unroll
                for (int j = 0; j < NB; j++) yy = fma(G[i * NB + j], v[j], yy);
                ao[(size_t)i * a.P] = g[h] - yy;
                c += v[i] * fma(0.5, yy, -g[h]);
            }
            a.cpart[((size_t)Y * R + r) * a.P + p] = c;
        }
    } else {
        a.cpart[((size_t)(2 * B) * R + r) * a.P + p] = 0.0;
        a.cpart[((size_t)(2 * B + 1) * R + r) * a.P + p] = 0.0;
    }
}

// per proposal t, both components of block B: v_n . a_o + 0.5 v_n G v_n
template <int L, int B>
__device__ __forceinline__ void lr_block_prop(const LR2Args& a, const double* __restrict__ gblk) {
    constexpr LRBlocks T = LRTab<L>::T;
    constexpr int NB = T.n[B];
    constexpr int NBMAX = lr_nbmax(L);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.T) return;
    if constexpr (NB > 0) {
        const int p = a.owner[t];
        const double4 xn = a.epos4[t];
        double vr[NB], vi[NB];
        lr_block_values2<L, B>(xn.x, xn.y, xn.z, xn.w, vr, vi);
#pragma unroll
        for (int comp = 0; comp < 2; comp++) {
            const double* __restrict__ G = gblk + T.off[B] + comp * NB * NB;   // wave-uniform: scalar loads
            const double* v = comp == 0 ? vr : vi;
            const int Y = 2 * B + comp;
            const double* ao = a.aold + (size_t)Y * NBMAX * a.P + p;
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < NB; i++) {
                double s = fma(0.5 * G[i * NB + i], v[i], ao[(size_t)i * a.P]);
#pragma unroll
                for (int j = 0; j < i; j++) s = fma(G[i * NB + j], v[j], s);
                acc = fma(v[i], s, acc);
            }
            a.lpart[(size_t)Y * a.T + t] = acc;
        }
    } else {
        a.lpart[(size_t)(2 * B) * a.T + t] = 0.0;
        a.lpart[(size_t)(2 * B + 1) * a.T + t] = 0.0;
    }
}

template <int L, int B>
__device__ __forceinline__ void lr_part_dispatch(int b, int r, const LR2Args& a, const double* gblk,
                                                 const double* g) {
    if constexpr (B < 6) {
        if (b == B) lr_block_part<L, B>(a, gblk, g, r);
        else lr_part_dispatch<L, B + 1>(b, r, a, gblk, g);
    }
}
template <int L, int B>
__device__ __forceinline__ void lr_prop_dispatch(int b, const LR2Args& a, const double* gblk) {
    if constexpr (B < 6) {
        if (b == B) lr_block_prop<L, B>(a, gblk);
        else lr_prop_dispatch<L, B + 1>(b, a, gblk);
    }
}

// a lane per moving charge / proposal does both components (real and imaginary parts)
// of its block; k_lr_part: grid.y = (block B of the operator (6)) x (row chunk, rg of
// them); k_lr_prop2: grid.y = block B
template <int L>
__global__ void __launch_bounds__(256) k_lr_part(LR2Args a, const double* __restrict__ gblk,
                                                 const double* __restrict__ gvec, int rg) {
    lr_part_dispatch<L, 0>((int)blockIdx.y / rg, (int)blockIdx.y % rg, a, gblk, gvec);
}

template <int L>
__global__ void __launch_bounds__(256) k_lr_prop2(LR2Args a, const double* __restrict__ gblk) {
    lr_prop_dispatch<L, 0>((int)blockIdx.y, a, gblk);
}
template <int L>
__global__ void __launch_bounds__(256) kb_lr_part(const LR2Args* __restrict__ args) {
    const LR2Args& a = args[blockIdx.z];
    lr_part_dispatch<L, 0>((int)blockIdx.y / LR_SPLIT_MAX, (int)blockIdx.y % LR_SPLIT_MAX, a, a.gblk, a.gvec);
}
template <int L>
__global__ void __launch_bounds__(256) kb_lr_prop2(const LR2Args* __restrict__ args) {
    const LR2Args& a = args[blockIdx.z];
    lr_prop_dispatch<L, 0>((int)blockIdx.y, a, a.gblk);
}


// level-0 multipole of a particle range: per block partial sums (deterministic),
// lane m of each wave owns column m.
template <int L>
__global__ void __launch_bounds__(64) k_multipole_partial(int n, int chunk, const double4* pos4, double* partial) {
    constexpr int NR = 2 * nhalf(L);
    const int b = blockIdx.x;
    const int m = threadIdx.x;
    double* dst = partial + (size_t)b * NR;
    if (m >= L) return;
    for (int l = m; l < L; l++) {
        dst[2 * hidx(l, m)] = 0.0;
        dst[2 * hidx(l, m) + 1] = 0.0;
    }
    const int i0 = b * chunk, i1 = min(n, i0 + chunk);
    for (int i = i0; i < i1; i++) {
        const double4 p = pos4[i];
        multipole_column<L>(m, p.x, p.y, p.z, p.w, dst, 1.0, true);
    }
}

// ============================================================================
// accept: move record, cell tables
// ============================================================================

// kmc_local.py:180-271 (extract: swap the last occupant into the freed slot) then
// kmc_local.py:274-359 (inject: append), then the particle arrays.
__device__ void store_update(const MoveRec* rec, int* occ, double4* slot_pos, int32_t* slot_id, int occmax,
                             double4* ppos4, int32_t* pcell, int* flags) {
    const int pid = rec->pid;
    const int oc = rec->old_cell, nc = rec->new_cell;
    const size_t ob = (size_t)oc * occmax, nb = (size_t)nc * occmax;
    const int oo = occ[oc];
    int index = -1, found = 0;
    for (int s = 0; s < oo; s++)
        if (slot_id[ob + s] == pid) { index = s; found++; }
    if (found != 1) { atomicOr(flags, 1); return; }
    slot_id[ob + index] = slot_id[ob + oo - 1];
    slot_pos[ob + index] = slot_pos[ob + oo - 1];
    occ[oc] = oo - 1;
    const int no = occ[nc];
    if (no >= occmax) { atomicOr(flags, 2); return; }
    slot_id[nb + no] = pid;
    slot_pos[nb + no] = make_double4(rec->new_pos[0], rec->new_pos[1], rec->new_pos[2], rec->q);
    occ[nc] = no + 1;
    ppos4[pid] = make_double4(rec->new_pos[0], rec->new_pos[1], rec->new_pos[2], rec->q);
    pcell[pid] = nc;
}

struct TailArgs {
    const MoveRec* rec;
    int* occ;
    double4* slot_pos;
    int32_t* slot_id;
    int occmax;
    double4* ppos4;
    int32_t* pcell;
    int* flags;
    int lr;                      // update M, g = G M and E_lr
    double* M;
    const double* gdat;
    const int64_t* gptr;
    const int32_t* gidx;
    double* g;
    double* elr;
    int64_t* counter;            // step counter to advance, or null
};

// everything of an accept except the leaf sweep, in one block: cell tables and
// particle arrays (thread 0), M += m(old, -q) then += m(new, +q)
// (kmc_full_long_range.py:270-316), g = G M, E_lr = 0.5 M.g, step counter.
template <int L>
__device__ __forceinline__ void accept_tail_body(const TailArgs& a) {
    constexpr int NR = 2 * nhalf(L);
    __shared__ double dold[NR], dnew[NR];
    __shared__ double part[256];
    const int tid = threadIdx.x;
    const bool valid = a.rec->valid != 0;
    if (valid && tid == 0) store_update(a.rec, a.occ, a.slot_pos, a.slot_id, a.occmax, a.ppos4, a.pcell, a.flags);
    if (valid && a.lr) {
        const MoveRec* r = a.rec;
        if (tid < L) {
            multipole_column<L>(tid, r->old_pos[0], r->old_pos[1], r->old_pos[2], r->q, dold, -1.0, false);
            multipole_column<L>(tid, r->new_pos[0], r->new_pos[1], r->new_pos[2], r->q, dnew, 1.0, false);
        }
        __syncthreads();
        for (int k = tid; k < NR; k += blockDim.x) a.M[k] = (a.M[k] + dold[k]) + dnew[k];
        __syncthreads();
        double acc = 0.0;
        for (int row = tid; row < NR; row += blockDim.x) {
            double y = 0.0;
            for (int64_t k = a.gptr[row]; k < a.gptr[row + 1]; k++) y = fma(a.gdat[k], a.M[a.gidx[k]], y);
            a.g[row] = y;
            acc = fma(a.M[row], y, acc);
        }
        part[tid] = acc;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if (tid < o) part[tid] += part[tid + o];
            __syncthreads();
        }
        if (tid == 0) *a.elr = 0.5 * part[0];
    }
    if (tid == 0 && a.counter) *a.counter += 1;
}
template <int L>
__global__ void __launch_bounds__(256) k_accept_tail(TailArgs a) { accept_tail_body<L>(a); }
template <int L>
__global__ void __launch_bounds__(256) kb_accept_tail(const TailArgs* __restrict__ args) {
    accept_tail_body<L>(args[blockIdx.z]);
}


// ============================================================================
// Metropolis selection on the device (one block)
// ============================================================================
// row of a cyclic per-step table for a step counter: counter mod nsteps in [0, nsteps)
// for every counter value, negative ones included
__device__ __forceinline__ int64_t step_of(int64_t counter, int64_t nsteps) {
    const int64_t r = counter % nsteps;
    return r < 0 ? r + nsteps : r;
}

struct SelectArgs {
    Geom g;
    int T;
    int n;
    int64_t nsteps;
    double kT;
    const double* dE;
    const double* rates;
    const double* bsum;
    const double* uniform;
    const int64_t* counter;
    int64_t* selected;
    const double4* epos4;
    const int32_t* ecell;
    const int32_t* owner;
    const int64_t* pid;
    const double4* ppos4;
    const int32_t* pcell;
    const int* occ;
    int occmax;
    double* energy;
    MoveRec* rec;
    int* flags;
    int* err;                    // bad-move count of this step's set-up (k_setup); reset here
};


// Rates and their sums over blocks of 64 proposals come from k_combine.  Each thread
// sums a contiguous chunk of block sums; chunk sums are scanned within each wave
// (shuffles) and across the 16 waves (LDS); the thread whose chunk holds u * total
// walks its block sums, then that block's rates, to find the proposal.
__device__ __forceinline__ void select_body(const SelectArgs& a) {
    __shared__ double wsum[16];
    __shared__ int pick;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int nt = blockDim.x;
    const int nw = nt >> 6;
    const int nb = (a.T + 63) / 64;
    const int chunk = (nb + nt - 1) / nt;
    const int b0 = min(nb, tid * chunk), b1 = min(nb, b0 + chunk);
    double s = 0.0;
    for (int b = b0; b < b1; b += 4) {
        double v4[4];
        load_parts(v4, a.bsum, b, b1, 1, 0);
        for (int j = 0; j < 4; j++)
            if (b + j < b1) s += v4[j];
    }
    if (tid == 0) pick = -1;
    double v = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    if (wid == 0) {
        double w = lane < nw ? wsum[lane] : 0.0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const double y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (lane < nw) wsum[lane] = w;
    }
    __syncthreads();
    // chunk intervals [lo, incl): each lower bound is the very value its predecessor
    // computed as upper bound, so the intervals tile [0, total) without gaps
    const double incl = v + (wid > 0 ? wsum[wid - 1] : 0.0);
    __shared__ double wlast[16];
    if (lane == 63) wlast[wid] = incl;
    const double prev = __shfl_up(incl, 1, 64);
    __syncthreads();
    const double lo = lane > 0 ? prev : (wid > 0 ? wlast[wid - 1] : 0.0);
    const int64_t step = step_of(*a.counter, a.nsteps);
    const double total = wlast[nw - 1];
    const double target = a.uniform[step] * total;
    if (target >= lo && target < incl && b1 > b0) {
        // the same running sum as one add per block sum, then one per rate, with the
        // loads issued 4 block sums / 8 rates at a time
        double c = lo;
        int k = -1;
        for (int g = b0; g < b1 && k < 0; g += 4) {
            double v4[4];
            load_parts(v4, a.bsum, g, b1, 1, 0);
            for (int j = 0; j < 4 && k < 0; j++) {
                const int b = g + j;
                if (b >= b1) break;
                const double cb = c + v4[j];
                if (target < cb || b == b1 - 1) {
                    const int t0 = b * 64, t1 = min(a.T, t0 + 64);
                    for (int u = t0; u < t1 && k < 0; u += 8) {
                        double r8[8];
                        load_parts(r8, a.rates, u, t1, 1, 0);
                        for (int i = 0; i < 8; i++) {
                            if (u + i >= t1) break;
                            c += r8[i];
                            if (target < c) { k = u + i; break; }
                        }
                    }
                    if (k < 0) k = t1 - 1;
                } else {
                    c = cb;
                }
            }
        }
        pick = k;
    }
    __syncthreads();
    if (tid == 0) {
        // a proposal outside the domain or beyond max_move (or a malformed es / pid):
        // nothing is accepted, flag 32 (the host paths raise KMC_E_BAD_MOVE instead)
        const int bad = *a.err;
        *a.err = 0;
        if (bad) {
            a.selected[step] = -1;
            a.rec->valid = 0;
            atomicOr(a.flags, 32);
            return;
        }
        int k = pick;
        if (k < 0) {   // target == total (u == 1) or no positive rate: last positive rate
            for (int b = nb - 1; b >= 0 && k < 0; b--) {
                if (!(a.bsum[b] > 0.0)) continue;
                for (int t = min(a.T, b * 64 + 64) - 1; t >= b * 64; t--)
                    if (a.rates[t] > 0.0) { k = t; break; }
            }
        }
        MoveRec* r = a.rec;
        a.selected[step] = k;
        if (k < 0) { r->valid = 0; atomicOr(a.flags, 4); return; }
        // every load before the first store to the record (the compiler cannot move a
        // load above a store that may alias it)
        const int p = a.owner[k];
        const double4 np = a.epos4[k];
        const int new_cell = a.ecell[k];
        const double dE = a.dE[k];
        const int pid = (int)a.pid[p];
        if (pid < 0 || pid >= a.n) { r->valid = 0; atomicOr(a.flags, 8); return; }
        const double4 op = a.ppos4[pid];
        const int old_cell = a.pcell[pid];
        const int occ_new = a.occ[new_cell];
        r->old_pos[0] = op.x; r->old_pos[1] = op.y; r->old_pos[2] = op.z;
        r->new_pos[0] = np.x; r->new_pos[1] = np.y; r->new_pos[2] = np.z;
        r->q = op.w;
        r->pid = pid;
        r->old_cell = old_cell;
        r->new_cell = new_cell;
        r->dE = dE;
        const bool full = (new_cell != old_cell) && (occ_new >= a.occmax);
        if (full) { r->valid = 0; atomicOr(a.flags, 2); return; }
        r->valid = 1;
        *a.energy += dE;
    }
}



// ============================================================================
// propose_with_dats set-up (mask compaction, kmc_mpi_decomp.py:548-614)
// ============================================================================



// synthetic move stream (benchmarks/tests)
struct MatArgs {
    Geom g;
    int n;
    int64_t nsteps;
    int P, k;
    const int64_t* pids_all;
    const double* disp_all;
    const int64_t* counter;
    const double4* ppos4;
    int64_t* pid;
    int64_t* es;
    double* new_pos;
    int* flags;
};
__device__ __forceinline__ void materialize_body(const MatArgs& a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = a.P, k = a.k;
    const int T = P * k;
    const int64_t step = step_of(*a.counter, a.nsteps);
    if (t <= P && t >= 0 && t < P + 1) a.es[t] = (int64_t)t * k;
    if (t >= T) return;
    const int p = t / k;
    int64_t i = a.pids_all[step * P + p];
    if (i < 0 || i >= a.n) { atomicOr(a.flags, 16); i = 0; }
    if (t % k == 0) a.pid[p] = i;
    const double4 o = a.ppos4[i];
    const double* d = a.disp_all + (step * T + t) * 3;
    double v[3] = {o.x + d[0], o.y + d[1], o.z + d[2]};
    const double h = 0.5 * a.g.E;
    for (int c = 0; c < 3; c++) {
        if (a.g.bc == KMC_BC_FREE_SPACE) {
            v[c] = v[c] < -h ? -h : (v[c] > h ? h : v[c]);
        } else {
            if (v[c] < -h) v[c] += a.g.E;
            if (v[c] >= h) v[c] -= a.g.E;
        }
        a.new_pos[3 * t + c] = v[c];
    }
}



// ============================================================================
// initial solve, hierarchical (large systems; coulomb_kmc/kmc_fmm.py:463-476,
// 956-959 take the initial leaf expansions from ppmd's FMM)
// ============================================================================
// Working variables are scaled Racah harmonics (m >= 0 stored, packed index
// l(l+1)/2 + m; the m < 0 entries are (-1)^m conj of the m > 0 ones):
//   R~_l^m(x) = R_l^m(x) / N_lm,  I~_l^m(x) = I_l^m(x) N_lm,  N_lm = sqrt((l-m)!(l+m)!)
// (R_l^m = r^l Y_l^m, I_l^m = r^{-l-1} Y_l^m as in the rest of this file).  At a level
// with cell width w:
//   multipole  Qh_lm = sum_i q_i conj(R~_l^m((x_i - c) / w))
//   local      Lh_jk = w^{j+1} N_jk L_jk   (L_jk the leaf-expansion layout)
// so the translations do not depend on the level:
//   M2M  Qh_p[j,k] = 2^-j sum Qh_c[l,m] conj(R~_{j-l}^{k-m}(b)),  b = (child - parent) / w_child
//   M2L  Lh_t[j,k] = (-1)^k sum (-1)^l I~_{j+l}^{m-k}(o) Qh_s[l,m],  o = (c_s - c_t) / w
//   L2L  Lh_c[l,m] = sum_{j>=l} 2^-(j+1) Lh_p[j,k] R~_{j-l}^{k-m}(b)
// The two finest levels' interaction lists (the 12^3 leaves around the leaf's
// grandparent block, minus the 27 near leaves) are added by direct P2L of their
// charges; M2L runs on levels 1 (2 for free space) .. R-2 only.

__host__ __device__ inline double rec_alpha(int l, int m) { return (2.0 * l - 1.0) / sqrt((double)(l * l - m * m)); }
__host__ __device__ inline double rec_beta(int l, int m) {
    return sqrt((double)((l - 1) * (l - 1) - m * m) / (double)(l * l - m * m));
}
__host__ __device__ inline void half_lm(int h, int& l, int& m) {
    l = 0;
    while ((l + 1) * (l + 2) / 2 <= h) l++;
    m = h - l * (l + 1) / 2;
}

// regular (IRR = false) or irregular (IRR = true) Schmidt solid harmonic of one (l, m), m >= 0
template <bool IRR>
__host__ __device__ inline Cplx solid_lm(int l, int m, double x, double y, double z) {
    const double r2 = x * x + y * y + z * z;
    double s = 1.0;
    for (int k = 1; k <= m; k++) s *= -sqrt((2.0 * k - 1.0) / (2.0 * k));
    Cplx w = {1.0, 0.0};
    for (int k = 0; k < m; k++) w = Cplx{w.re * x - w.im * y, w.re * y + w.im * x};
    double b1 = s, b2 = 0.0;
    if (IRR) {
        const double ir2 = 1.0 / r2;
        double p = 1.0 / sqrt(r2);
        for (int k = 0; k < m; k++) p *= ir2;
        b1 = s * p;
        for (int k = m + 1; k <= l; k++) {
            const double b = (rec_alpha(k, m) * z * b1 - rec_beta(k, m) * b2) * ir2;
            b2 = b1;
            b1 = b;
        }
    } else {
        for (int k = m + 1; k <= l; k++) {
            const double b = rec_alpha(k, m) * z * b1 - rec_beta(k, m) * r2 * b2;
            b2 = b1;
            b1 = b;
        }
    }
    return Cplx{b1 * w.re, b1 * w.im};
}

__device__ __forceinline__ double2 conj_sgn(double2 v, int m) {   // (-1)^m conj(v)
    return (m & 1) ? make_double2(-v.x, v.y) : make_double2(v.x, -v.y);
}
// entry (l, m) of a packed half expansion for any -l <= m <= l
__device__ __forceinline__ double2 half_get(const double2* a, int l, int m) {
    const double2 v = a[hidx(l, m < 0 ? -m : m)];
    return m < 0 ? conj_sgn(v, -m) : v;
}

struct FmmLevel {
    int n;                 // cells per side
    double w;              // cell width
};




// M2L at one level: block per target cell, lane per packed output (j, k).  For each
// source of the target's interaction list (children of the parent's neighbours not
// adjacent to the target; periodic: indices wrap, the displacement does not) the
// block builds the I~(o) table in LDS and stages the source multipole.
constexpr int FMM_M2L_THREADS = 256;

// Leaf P2L of the charges in the 12^3 leaves around the leaf's grandparent block
// (offsets o = -4-b .. 7-b per axis, b = leaf index mod 4) except its 27 near leaves.
// One wave per 64 leaves sharing b; grid.y = column-pair group (as the accept sweep).
template <int L, int G, int NP>
__device__ __forceinline__ void p2l_near_group(const SweepArgs& a, const int* occ, const double4* slot_pos, int occmax,
                                               int cell, int bx, int by, int bz) {
    constexpr int P0 = G * NP;
    constexpr int W = L + 1;
    double ar[NP][W], ai[NP][W];
#pragma unroll
    for (int jj = 0; jj < NP; jj++)
#pragma unroll
        for (int kk = 0; kk < W; kk++) { ar[jj][kk] = 0.0; ai[jj][kk] = 0.0; }
    const int n = a.g.ncps;
    const double E = a.g.E;
    const bool periodic = a.g.bc != KMC_BC_FREE_SPACE;
    const int tx = cell % n, ty = (cell / n) % n, tz = cell / (n * n);
    const double cx = a.g.c0 + tx * a.g.w, cy = a.g.c0 + ty * a.g.w, cz = a.g.c0 + tz * a.g.w;
#pragma unroll 1
    for (int oz = -4 - bz; oz <= 7 - bz; oz++) {
        int sz = tz + oz;
        double shz = 0.0;
        if (sz < 0) { sz += n; shz = -E; } else if (sz >= n) { sz -= n; shz = E; }
        if (!periodic && shz != 0.0) continue;
#pragma unroll 1
        for (int oy = -4 - by; oy <= 7 - by; oy++) {
            int sy = ty + oy;
            double shy = 0.0;
            if (sy < 0) { sy += n; shy = -E; } else if (sy >= n) { sy -= n; shy = E; }
            if (!periodic && shy != 0.0) continue;
#pragma unroll 1
            for (int ox = -4 - bx; ox <= 7 - bx; ox++) {
                if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
                int sx = tx + ox;
                double shx = 0.0;
                if (sx < 0) { sx += n; shx = -E; } else if (sx >= n) { sx -= n; shx = E; }
                if (!periodic && shx != 0.0) continue;
                const int s = sx + n * (sy + n * sz);
                const int o = occ[s];
#pragma unroll 1
                for (int q = 0; q < o; q++) {
                    const double4 p = slot_pos[(size_t)s * occmax + q];
                    const double dx[1] = {(p.x + shx) - cx};
                    const double qe[1] = {p.w};
                    sweep_images<L, P0, NP, 1>(dx, (p.y + shy) - cy, (p.z + shz) - cz, qe, ar, ai);
                }
            }
        }
    }
    sweep_group_store<L, P0, NP>(a.lexp + (size_t)cell * nhalf(L), ar, ai);
}
template <int L, int NP, int G0>
__device__ __forceinline__ void p2l_near_dispatch(const SweepArgs& a, const int* occ, const double4* slot_pos, int occmax,
                                                  int cell, int group, int bx, int by, int bz) {
    constexpr int NG = ((L + 1) / 2 + NP - 1) / NP;
    if constexpr (G0 < NG) {
        if (group == G0) p2l_near_group<L, G0, NP>(a, occ, slot_pos, occmax, cell, bx, by, bz);
        else p2l_near_dispatch<L, NP, G0 + 1>(a, occ, slot_pos, occmax, cell, group, bx, by, bz);
    }
}
// grid.x = 64 (b classes) x ceil(ng^3 / 64) chunks, ng = ncps / 4; block = 64
template <int L>
__global__ void __launch_bounds__(64) k_fmm_p2l_near(SweepArgs a, const int* occ, const double4* slot_pos, int occmax) {
    const int b = blockIdx.x % 64;
    const int chunk = blockIdx.x / 64;
    const int bx = b & 3, by = (b >> 2) & 3, bz = b >> 4;
    const int ng = a.g.ncps / 4;
    const int gi = chunk * 64 + threadIdx.x;
    if (gi >= ng * ng * ng) return;
    const int gx = gi % ng, gy = (gi / ng) % ng, gz = gi / (ng * ng);
    const int cell = (4 * gx + bx) + a.g.ncps * ((4 * gy + by) + a.g.ncps * (4 * gz + bz));
    p2l_near_dispatch<L, 1, 0>(a, occ, slot_pos, occmax, cell, (int)blockIdx.y, bx, by, bz);
}

// ============================================================================
// potential at points (KMCFMM.eval_field, coulomb_kmc/kmc_fmm.py:544-567)
// ============================================================================



// ============================================================================
// charge energies, energies of small groups, injection and extraction of charges
// (coulomb_kmc/kmc_inject_extract.py:82-424)
// ============================================================================



// column m of the packed multipole of charge p.w at p about the origin, added to Mb
// (the vector v(x) of k_eval_lr scaled by the charge: M = sum_i q_i v(x_i))
__device__ void multipole_column_add(int L, int m, double4 p, double* Mb) {
    const double r2 = p.x * p.x + p.y * p.y + p.z * p.z;
    double s = 1.0;
    Cplx C = {1.0, 0.0};
    for (int k = 1; k <= m; k++) {
        s *= -sqrt((2.0 * k - 1.0) / (2.0 * k));
        C = Cplx{C.re * p.x + C.im * p.y, C.im * p.x - C.re * p.y};
    }
    double b2 = 0.0, b1 = s;
    for (int l = m; l < L; l++) {
        if (l > m) {
            const double b = rec_alpha(l, m) * p.z * b1 - rec_beta(l, m) * r2 * b2;
            b2 = b1;
            b1 = b;
        }
        const int h = hidx(l, m);
        Mb[2 * h] = fma(p.w * b1, C.re, Mb[2 * h]);
        Mb[2 * h + 1] = fma(p.w * b1, C.im, Mb[2 * h + 1]);
    }
}










// ============================================================================
// lattice bookkeeping on the device (Algorithm 2 of the reference's examples:
// docs/source/example_bookkeeping.rst, examples/propose_with_dats_example.py:217-312)
// ============================================================================



// ============================================================================
// order dispatch tables
// ============================================================================
typedef void (*SweepLaunch)(SweepArgs, int, hipStream_t, int, int, int);
typedef void (*EvalLaunch)(EvalArgs, hipStream_t);
typedef void (*LRLaunch)(LR2Args, hipStream_t);
typedef void (*MpLaunch)(int, int, const double4*, double*, int, hipStream_t);
typedef void (*MdLaunch)(TailArgs, hipStream_t);
typedef void (*P2LNearLaunch)(SweepArgs, const int*, const double4*, int, hipStream_t);

template <int L>
static void launch_p2l_near(SweepArgs a, const int* occ, const double4* slot_pos, int occmax, hipStream_t s) {
    const int ng = a.g.ncps / 4;
    const int chunks = (ng * ng * ng + 63) / 64;
    hipLaunchKernelGGL((k_fmm_p2l_near<L>), dim3(64 * chunks, (L + 1) / 2), dim3(64), 0, s, a, occ, slot_pos, occmax);
}

template <int L, int NP>
static void launch_sweep_np(SweepArgs a, int ncells, hipStream_t s, int block, int split) {
    const int ng = ((L + 1) / 2 + NP - 1) / NP;
    if (block == 0 && ncells % 64 == 0 && 64 * ng <= 512) {
        if (a.g.bc == KMC_BC_FREE_SPACE)
            hipLaunchKernelGGL((k_sweep_fused<L, NP, false>), dim3(ncells / 64), dim3(64, ng), 0, s, a);
        else
            hipLaunchKernelGGL((k_sweep_fused<L, NP, true>), dim3(ncells / 64), dim3(64, ng), 0, s, a);
        return;
    }
    if (block == 0) block = 64;
    if (a.g.bc == KMC_BC_FREE_SPACE) {
        hipLaunchKernelGGL((k_sweep<L, NP, false, 1>), dim3((ncells + block - 1) / block, ng), dim3(block), 0, s, a);
    } else if (split == 3) {
        hipLaunchKernelGGL((k_sweep<L, NP, true, 3>), dim3((ncells + 63) / 64, ng), dim3(64, 3), 0, s, a);
    } else {
        hipLaunchKernelGGL((k_sweep<L, NP, true, 1>), dim3((ncells + block - 1) / block, ng), dim3(block), 0, s, a);
    }
}
// the blocked sweep: two column pairs per lane, one source image per recurrence pass
// (both sources' images together need more than 256 VGPRs)
template <int L>
static void launch_sweep_block(SweepArgs a, hipStream_t s) {
    constexpr int ng = ((L + 1) / 2 + 1) / 2;
    const int nb = a.g.ncps / 4;
    const dim3 grid(nb * nb * nb), block(64, ng);
    if (a.g.bc == KMC_BC_FREE_SPACE) hipLaunchKernelGGL((k_sweep_block<L, 2, 1, false>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_sweep_block<L, 2, 1, true>), grid, block, 0, s, a);
}
static bool sweep_block_fits(const SweepArgs& a, int L) {
    return a.mode == 0 && a.g.ncps % 4 == 0 && 64 * (((L + 1) / 2 + 1) / 2) <= 512;
}
template <int L>
static void launch_sweep(SweepArgs a, int ncells, hipStream_t s, int block, int np, int split) {
    // block 1: the blocked sweep with order truncation (the move record only; leaf grids
    // in multiples of 4 per side); other shapes otherwise
    if (block == 1 && sweep_block_fits(a, L)) {
        launch_sweep_block<L>(a, s);
        return;
    }
    if (block == 1) block = 0;
    if (np == 1) launch_sweep_np<L, 1>(a, ncells, s, block, split);
    else launch_sweep_np<L, 2>(a, ncells, s, block, split);
}

template <int L>
static void launch_far(EvalArgs a, hipStream_t s) {
    if (a.n <= 0) return;
    constexpr int npair = (L + 1) / 2;
    constexpr int py = far_pairs_per_group(L);   // pairs per workgroup, <= 8 waves
    hipLaunchKernelGGL(k_far<L>, dim3((a.n + 63) / 64, (npair + py - 1) / py), dim3(64, py), 0, s, a);
}
template <int L>
static void launch_lr(LR2Args a, hipStream_t s) {
    if (a.T <= 0 || a.P <= 0) return;
    hipLaunchKernelGGL(k_lr_part<L>, dim3((a.P + 255) / 256, 6 * a.rsplit), dim3(256), 0, s, a, a.gblk, a.gvec,
                       a.rsplit);
    hipLaunchKernelGGL(k_lr_prop2<L>, dim3((a.T + 255) / 256, 6), dim3(256), 0, s, a, a.gblk);
}
template <int L>
static void launch_mp(int n, int chunk, const double4* p, double* partial, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_multipole_partial<L>, dim3(nblk), dim3(64), 0, s, n, chunk, p, partial);
}
template <int L>
static void launch_md(TailArgs a, hipStream_t s) {
    hipLaunchKernelGGL(k_accept_tail<L>, dim3(1), dim3(256), 0, s, a);
}

// batched launches: argument tables in device memory, system = blockIdx.z
struct BatchShape {
    int n;           // systems in the table
    int max_ne;      // largest T + P
    int max_P, max_T;
    int max_cells;   // largest leaf count (a multiple of 64)
};
typedef void (*BFarLaunch)(const EvalArgs*, BatchShape, hipStream_t);
typedef void (*BLRLaunch)(const LR2Args*, BatchShape, hipStream_t);
typedef void (*BSweepLaunch)(const SweepArgs*, const int2*, int, BatchShape, bool, hipStream_t);
typedef void (*BTailLaunch)(const TailArgs*, BatchShape, hipStream_t);

template <int L>
static void blaunch_far(const EvalArgs* t, BatchShape b, hipStream_t s) {
    constexpr int npair = (L + 1) / 2;
    constexpr int py = far_pairs_per_group(L);
    hipLaunchKernelGGL(kb_far<L>, dim3((b.max_ne + 63) / 64, (npair + py - 1) / py, b.n), dim3(64, py), 0, s, t);
}
template <int L>
static void blaunch_lr(const LR2Args* t, BatchShape b, hipStream_t s) {
    hipLaunchKernelGGL(kb_lr_part<L>, dim3((b.max_P + 255) / 256, 6 * LR_SPLIT_MAX, b.n), dim3(256), 0, s, t);
    hipLaunchKernelGGL(kb_lr_prop2<L>, dim3((b.max_T + 255) / 256, 6, b.n), dim3(256), 0, s, t);
}
// accept sweep of a group: the blocked layout over the flat (system, block) table when
// one is given (nflat blocks), else the fused layout (grid sized for the largest system)
template <int L>
static void blaunch_sweep(const SweepArgs* t, const int2* tab, int nflat, BatchShape b, bool per, hipStream_t s) {
    constexpr int ng = ((L + 1) / 2 + 1) / 2;
    if (tab && nflat > 0 && 64 * ng <= 512) {
        if (per) hipLaunchKernelGGL((kb_sweep_block<L, 2, 1, true>), dim3(nflat), dim3(64, ng), 0, s, t, tab);
        else hipLaunchKernelGGL((kb_sweep_block<L, 2, 1, false>), dim3(nflat), dim3(64, ng), 0, s, t, tab);
        return;
    }
    if (per) hipLaunchKernelGGL((kb_sweep_fused<L, 2, true>), dim3(b.max_cells / 64, 1, b.n), dim3(64, ng), 0, s, t);
    else hipLaunchKernelGGL((kb_sweep_fused<L, 2, false>), dim3(b.max_cells / 64, 1, b.n), dim3(64, ng), 0, s, t);
}
template <int L>
static void blaunch_tail(const TailArgs* t, BatchShape b, hipStream_t s) {
    hipLaunchKernelGGL(kb_accept_tail<L>, dim3(1, 1, b.n), dim3(256), 0, s, t);
}

struct OrderFns {
    SweepLaunch sweep;
    EvalLaunch far;
    LRLaunch lr;
    MpLaunch mp;
    MdLaunch md;
    P2LNearLaunch p2ln;
    BFarLaunch bfar;
    BLRLaunch blr;
    BSweepLaunch bsweep;
    BTailLaunch btail;
};
template <int L>
static OrderFns fns_for() {
    return {launch_sweep<L>, launch_far<L>,  launch_lr<L>,    launch_mp<L>,     launch_md<L>,
            launch_p2l_near<L>, blaunch_far<L>, blaunch_lr<L>, blaunch_sweep<L>, blaunch_tail<L>};
}
