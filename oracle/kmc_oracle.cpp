// ORACLE - test infrastructure only.
//
// CPU FP64 restatement of the reference's FMM-KMC propose/accept path
// (ppmd/coulomb_kmc), used by tests/ as the checker and by bench.py's
// cpu_baseline leg as the timed CPU path.  Nothing in the product
// (coulomb_kmc_amd/) links, loads or calls this file.
//
// It follows the reference's own data structures and loop orders:
//   * the halo'd local particle store  [lsd^3][occmax][5] with periodically
//     shifted halo copies            coulomb_kmc/kmc_local.py:558-824
//   * direct_new / direct_old        coulomb_kmc/kmc_local.py:928-1031
//   * indirect_interactions          coulomb_kmc/kmc_octal.py:495-707
//   * self_interaction               coulomb_kmc/kmc_fmm_self_interaction.py:38-267
//   * long_range_energy              coulomb_kmc/kmc_full_long_range.py:142-194, 391-896
//   * diff_lib                       coulomb_kmc/kmc_fmm.py:583-663
//   * setup_move (mask compaction)   coulomb_kmc/kmc_mpi_decomp.py:347-629
//   * extract / inject               coulomb_kmc/kmc_local.py:180-370
//   * accept_local_exp (hop case)    coulomb_kmc/kmc_octal.py:710-1253
//   * long range accept              coulomb_kmc/kmc_full_long_range.py:270-316
//   * basis routines                 coulomb_kmc/kmc_expansion_tools.py:38-350
// with spherical coordinates from atan2 (coulomb_kmc/common.py:6-33) and the
// associated Legendre functions by the textbook three-term recurrence (the
// reference generates this code through ppmd's SphGen, which is not in the tree;
// its values are pinned by scipy.special.lpmv in tests/host/test_c_sph_harm.py).
//
// Build: oracle/Makefile  ->  oracle/libkmc_oracle.so

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

enum { BC_FREE = 0, BC_27 = 1, BC_PBC = 2, BC_FF = 3 };

// stencil order of coulomb_kmc/common.py:51-81 (x fastest)
static int OFF[27][3];
static void init_offsets() {
    int k = 0;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                OFF[k][0] = x; OFF[k][1] = y; OFF[k][2] = z; k++;
            }
}

struct Basis {
    int L;
    std::vector<double> H;   // H[l*L+m] = sqrt((l-m)!/(l+m)!)
    explicit Basis(int L_) : L(L_), H(L_ * L_, 0.0) {
        for (int l = 0; l < L; l++)
            for (int m = 0; m <= l; m++)
                H[l * L + m] = std::exp(0.5 * (std::lgamma(l - m + 1.0) - std::lgamma(l + m + 1.0)));
    }
    // Y_l^m, m >= 0, at (theta, phi):  yre/yim[l*L+m]
    void Y(double theta, double phi, double* yre, double* yim) const {
        const double ct = std::cos(theta), st = std::sin(theta);
        std::vector<double> P(L * L, 0.0);
        double pmm = 1.0;
        for (int m = 0; m < L; m++) {
            if (m > 0) pmm *= -(2.0 * m - 1.0) * st;
            P[m * L + m] = pmm;
            if (m + 1 < L) P[(m + 1) * L + m] = (2.0 * m + 1.0) * ct * pmm;
            for (int l = m + 2; l < L; l++)
                P[l * L + m] = ((2.0 * l - 1.0) * ct * P[(l - 1) * L + m] - (l + m - 1.0) * P[(l - 2) * L + m]) / (l - m);
        }
        for (int l = 0; l < L; l++)
            for (int m = 0; m <= l; m++) {
                const double v = H[l * L + m] * P[l * L + m];
                yre[l * L + m] = v * std::cos(m * phi);
                yim[l * L + m] = v * std::sin(m * phi);
            }
    }
};

static inline void spherical(double dx, double dy, double dz, double* r, double* th, double* ph) {
    const double xy = dx * dx + dy * dy;
    *r = std::sqrt(xy + dz * dz);
    *th = std::atan2(std::sqrt(xy), dz);
    *ph = std::atan2(dy, dx);
}

// kmc_expansion_tools.py:154-204 : q * Re sum L_lm r^l Y_l^m
static double local_eval(const Basis& B, const double* lexp, double r, double th, double ph) {
    const int L = B.L, L2 = L * L;
    std::vector<double> yre(L2), yim(L2);
    B.Y(th, ph, yre.data(), yim.data());
    double e = 0.0, rl = 1.0;
    for (int l = 0; l < L; l++) {
        for (int m = -l; m <= l; m++) {
            const int am = m < 0 ? -m : m;
            const double yr = yre[l * L + am];
            const double yi = m < 0 ? -yim[l * L + am] : yim[l * L + am];
            const int idx = l * l + l + m;
            e += rl * (lexp[idx] * yr - lexp[L2 + idx] * yi);
        }
        rl *= r;
    }
    return e;
}

// kmc_expansion_tools.py:211-226 : out_lm += q r^{-l-1} Y_l^{-m}
static void p2l(const Basis& B, double q, double r, double th, double ph, double* out) {
    const int L = B.L, L2 = L * L;
    std::vector<double> yre(L2), yim(L2);
    B.Y(th, ph, yre.data(), yim.data());
    const double ir = 1.0 / r;
    double rl = ir;
    for (int l = 0; l < L; l++) {
        for (int m = -l; m <= l; m++) {
            const int am = m < 0 ? -m : m;
            // Y_l^{-m} = conj(Y_l^m)
            const double yr = yre[l * L + am];
            const double yi = (-m) < 0 ? -yim[l * L + am] : yim[l * L + am];
            const int idx = l * l + l + m;
            out[idx] += yr * rl * q;
            out[L2 + idx] += yi * rl * q;
        }
        rl *= ir;
    }
}

// kmc_expansion_tools.py:45-58 (P2M) and :99-113 (dot vector); both return the
// same numbers (Y_l^{-m} = conj Y_l^m), computed here the reference's two ways.
static void p2m(const Basis& B, double q, double r, double th, double ph, double* out) {
    const int L = B.L, L2 = L * L;
    std::vector<double> yre(L2), yim(L2);
    B.Y(th, ph, yre.data(), yim.data());
    double rl = 1.0;
    for (int l = 0; l < L; l++) {
        for (int m = -l; m <= l; m++) {
            const int am = m < 0 ? -m : m;
            const double yr = yre[l * L + am];
            const double yi = (-m) < 0 ? -yim[l * L + am] : yim[l * L + am];
            const int idx = l * l + l + m;
            out[idx] += yr * rl * q;
            out[L2 + idx] += yi * rl * q;
        }
        rl *= r;
    }
}
static void dot_vec(const Basis& B, double q, double r, double th, double ph, double* out) {
    const int L = B.L, L2 = L * L;
    std::vector<double> yre(L2), yim(L2);
    B.Y(th, ph, yre.data(), yim.data());
    double rl = 1.0;
    for (int l = 0; l < L; l++) {
        for (int m = -l; m <= l; m++) {
            const int am = m < 0 ? -m : m;
            const double yr = yre[l * L + am];
            const double yi = m < 0 ? -yim[l * L + am] : yim[l * L + am];
            const int idx = l * l + l + m;
            out[idx] += yr * rl * q;
            out[L2 + idx] += -1.0 * yi * rl * q;
        }
        rl *= r;
    }
}

// ---------------------------------------------------------------------------
// The same basis routines on a real type R (double or long double): used by the
// checker below, which re-evaluates proposal energies of the stored state in
// extended precision.  Same formulas as Basis / local_eval / p2m / dot_vec above.
// ---------------------------------------------------------------------------
template <typename R>
struct BasisT {
    int L;
    std::vector<R> H;
    explicit BasisT(int L_) : L(L_), H(L_ * L_, R(0)) {
        for (int l = 0; l < L; l++)
            for (int m = 0; m <= l; m++)
                H[l * L + m] = std::exp(R(0.5) * (std::lgamma(R(l - m + 1)) - std::lgamma(R(l + m + 1))));
    }
    void Y(R theta, R phi, R* yre, R* yim) const {
        const R ct = std::cos(theta), st = std::sin(theta);
        std::vector<R> P(L * L, R(0));
        R pmm = 1;
        for (int m = 0; m < L; m++) {
            if (m > 0) pmm *= -(R(2) * m - 1) * st;
            P[m * L + m] = pmm;
            if (m + 1 < L) P[(m + 1) * L + m] = (R(2) * m + 1) * ct * pmm;
            for (int l = m + 2; l < L; l++)
                P[l * L + m] = ((R(2) * l - 1) * ct * P[(l - 1) * L + m] - (R(l + m) - 1) * P[(l - 2) * L + m]) / R(l - m);
        }
        for (int l = 0; l < L; l++)
            for (int m = 0; m <= l; m++) {
                const R v = H[l * L + m] * P[l * L + m];
                yre[l * L + m] = v * std::cos(R(m) * phi);
                yim[l * L + m] = v * std::sin(R(m) * phi);
            }
    }
};
template <typename R>
static inline void spherical_t(R dx, R dy, R dz, R* r, R* th, R* ph) {
    const R xy = dx * dx + dy * dy;
    *r = std::sqrt(xy + dz * dz);
    *th = std::atan2(std::sqrt(xy), dz);
    *ph = std::atan2(dy, dx);
}
template <typename R>
static R local_eval_t(const BasisT<R>& B, const double* lexp, R r, R th, R ph) {
    const int L = B.L, L2 = L * L;
    std::vector<R> yre(L2), yim(L2);
    B.Y(th, ph, yre.data(), yim.data());
    R e = 0, rl = 1;
    for (int l = 0; l < L; l++) {
        for (int m = -l; m <= l; m++) {
            const int am = m < 0 ? -m : m;
            const R yr = yre[l * L + am];
            const R yi = m < 0 ? -yim[l * L + am] : yim[l * L + am];
            const int idx = l * l + l + m;
            e += rl * (R(lexp[idx]) * yr - R(lexp[L2 + idx]) * yi);
        }
        rl *= r;
    }
    return e;
}
// P2M (sign of the imaginary part +) or dot vector (-) of charge q at (r, th, ph) into out
template <typename R>
static void p2m_t(const BasisT<R>& B, R q, R r, R th, R ph, R* out, bool dotvec) {
    const int L = B.L, L2 = L * L;
    std::vector<R> yre(L2), yim(L2);
    B.Y(th, ph, yre.data(), yim.data());
    R rl = 1;
    for (int l = 0; l < L; l++) {
        for (int m = -l; m <= l; m++) {
            const int am = m < 0 ? -m : m;
            const R yr = yre[l * L + am];
            const R yi = m < 0 ? -yim[l * L + am] : yim[l * L + am];
            const int idx = l * l + l + m;
            out[idx] += yr * rl * q;
            out[L2 + idx] += (dotvec ? -yi : (m > 0 ? -yim[l * L + am] : yim[l * L + am])) * rl * q;
        }
        rl *= r;
    }
}

struct Oracle {
    int64_t N;
    double E, max_move, eu;
    int R, L, bc, ncps, pad, lsd;
    int occmax;
    Basis B;
    double w;
    // particles (local index == global id, single rank)
    std::vector<double> pos, q;
    std::vector<int64_t> cell;        // linear primary cell
    // halo'd store
    std::vector<int64_t> cell_ind;    // [lsd] primary index per local index
    std::vector<int64_t> pfac;        // [lsd] periodic factor
    std::vector<int64_t> occ;         // [lsd^3]
    std::vector<double> pdata;        // [lsd^3][occmax][5]
    std::vector<int64_t> pids;        // [lsd^3][occmax]
    // expansions on the halo'd store and centres
    std::vector<double> lexp;         // [lsd^3][2L^2]
    std::vector<double> centres;      // [lsd^3][3]   shifted centres
    std::vector<double> ocentres;     // [lsd^3][3]   primary centres
    // long range
    std::vector<double> M, Ev;        // [2L^2]
    std::vector<double> lr_data;
    std::vector<int64_t> lr_indptr, lr_indices;
    double dip[3] = {0, 0, 0};
    bool have_lr = false;
    double self27 = 0.0;
    double energy = 0.0;

    Oracle(int64_t n, double e, int r, int l, int b, double mm, double eunit)
        : N(n), E(e), max_move(mm), eu(eunit), R(r), L(l), bc(b), B(l) {
        ncps = 1 << (R - 1);
        w = E / ncps;
        pad = 2 + (int)std::ceil(max_move / w);
        lsd = ncps + 2 * pad;
        cell_ind.resize(lsd);
        pfac.resize(lsd);
        for (int k = 0; k < lsd; k++) {
            const int64_t g = k - pad;
            int64_t f = g >= 0 ? g / ncps : -((-g + ncps - 1) / ncps);
            pfac[k] = f;
            cell_ind[k] = g - f * ncps;
        }
        occmax = 0;
        // kmc_fmm_self_interaction.py:48-54
        for (int k = 0; k < 27; k++) {
            if (OFF[k][0] == 0 && OFF[k][1] == 0 && OFF[k][2] == 0) continue;
            const double dx = E * OFF[k][0], dy = E * OFF[k][1], dz = E * OFF[k][2];
            self27 -= 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        M.assign(2 * L * L, 0.0);
        Ev.assign(2 * L * L, 0.0);
    }

    int64_t lsd3() const { return (int64_t)lsd * lsd * lsd; }
    int64_t lidx(int kx, int ky, int kz) const { return kx + (int64_t)lsd * (ky + (int64_t)lsd * kz); }

    // kmc_fmm.py:229-242 (cell of a position, divide form, truncate and clamp)
    void bin(const double* p, int64_t* c3) const {
        for (int d = 0; d < 3; d++) {
            int64_t c = (int64_t)((0.5 * E + p[d]) / w);
            if (c < 0) c = 0;
            if (c > ncps - 1) c = ncps - 1;
            c3[d] = c;
        }
    }
    int64_t lin(const int64_t* c3) const { return c3[0] + ncps * (c3[1] + ncps * c3[2]); }
    void unlin(int64_t c, int64_t* c3) const {
        c3[0] = c % ncps; c3[1] = (c / ncps) % ncps; c3[2] = c / ((int64_t)ncps * ncps);
    }
    // kmc_local.py:796-817 : which halo copies hold particles
    bool copy_allowed(int kx, int ky, int kz) const {
        const int64_t f[3] = {pfac[kx], pfac[ky], pfac[kz]};
        for (int d = 0; d < 3; d++) {
            if (bc == BC_FREE && f[d] != 0) return false;
            if (bc == BC_27 && (f[d] > 1 || f[d] < -1)) return false;
        }
        return true;
    }

    void set_particles(const double* p, const double* qq) {
        pos.assign(p, p + 3 * N);
        q.assign(qq, qq + N);
        cell.resize(N);
        std::vector<int64_t> cnt((size_t)ncps * ncps * ncps, 0);
        for (int64_t i = 0; i < N; i++) {
            int64_t c3[3];
            bin(&pos[3 * i], c3);
            cell[i] = lin(c3);
            cnt[cell[i]]++;
        }
        int64_t mx = 1;
        for (auto c : cnt) mx = std::max(mx, c);
        occmax = (int)mx;
        build_store();
        // centres, kmc_octal.py:397-428
        centres.assign(lsd3() * 3, 0.0);
        ocentres.assign(lsd3() * 3, 0.0);
        for (int kz = 0; kz < lsd; kz++)
            for (int ky = 0; ky < lsd; ky++)
                for (int kx = 0; kx < lsd; kx++) {
                    const int64_t li = lidx(kx, ky, kz);
                    const int k3[3] = {kx, ky, kz};
                    for (int d = 0; d < 3; d++) {
                        const double start = -0.5 * E + 0.5 * w;
                        centres[3 * li + d] = start + (double)(k3[d] - pad) * w;
                        ocentres[3 * li + d] = start + (double)cell_ind[k3[d]] * w;
                    }
                }
        lexp.assign(lsd3() * 2 * L * L, 0.0);
    }

    // kmc_local.py:582-824 : particles of a cell in ascending local index; halo copies shifted
    void build_store() {
        const int64_t n3 = lsd3();
        occ.assign(n3, 0);
        pdata.assign(n3 * occmax * 5, 0.0);
        pids.assign(n3 * occmax, 0);
        std::vector<std::vector<int64_t>> members((size_t)ncps * ncps * ncps);
        for (int64_t i = 0; i < N; i++) members[cell[i]].push_back(i);
        for (int kz = 0; kz < lsd; kz++)
            for (int ky = 0; ky < lsd; ky++)
                for (int kx = 0; kx < lsd; kx++) {
                    const int64_t li = lidx(kx, ky, kz);
                    if (!copy_allowed(kx, ky, kz)) { occ[li] = 0; continue; }
                    const int64_t c3[3] = {cell_ind[kx], cell_ind[ky], cell_ind[kz]};
                    const auto& mem = members[lin(c3)];
                    occ[li] = (int64_t)mem.size();
                    for (size_t s = 0; s < mem.size(); s++) {
                        const int64_t i = mem[s];
                        double* d = &pdata[(li * occmax + s) * 5];
                        d[0] = pos[3 * i + 0] + pfac[kx] * E;
                        d[1] = pos[3 * i + 1] + pfac[ky] * E;
                        d[2] = pos[3 * i + 2] + pfac[kz] * E;
                        d[3] = q[i];
                        std::memcpy(&d[4], &i, sizeof(int64_t));
                        pids[li * occmax + s] = i;
                    }
                }
    }

    void resize_store(int newmax) {
        if (newmax <= occmax) return;
        const int64_t n3 = lsd3();
        std::vector<double> nd(n3 * newmax * 5, 0.0);
        std::vector<int64_t> ni(n3 * newmax, 0);
        for (int64_t c = 0; c < n3; c++)
            for (int s = 0; s < occmax; s++) {
                for (int k = 0; k < 5; k++) nd[(c * newmax + s) * 5 + k] = pdata[(c * occmax + s) * 5 + k];
                ni[c * newmax + s] = pids[c * occmax + s];
            }
        pdata.swap(nd);
        pids.swap(ni);
        occmax = newmax;
    }

    // local store index of a primary cell (single rank: offset by pad)
    int64_t store_cell(int64_t c) const {
        int64_t c3[3];
        unlin(c, c3);
        return lidx((int)(c3[0] + pad), (int)(c3[1] + pad), (int)(c3[2] + pad));
    }

    // ---------------------------------------------------------------- near field
    // kmc_local.py:956-1019
    double direct(const double* p, double qq, int64_t lcell, int64_t exclude_id) const {
        double e = 0.0;
        for (int k = 0; k < 27; k++) {
            const int64_t jc = lcell + OFF[k][0] + (int64_t)lsd * (OFF[k][1] + (int64_t)lsd * OFF[k][2]);
            const int64_t n = occ[jc];
            for (int64_t j = 0; j < n; j++) {
                const double* d = &pdata[(jc * occmax + j) * 5];
                const double dx = p[0] - d[0], dy = p[1] - d[1], dz = p[2] - d[2];
                const double contrib = d[3] / std::sqrt(dx * dx + dy * dy + dz * dz);
                if (exclude_id >= 0) e += (pids[jc * occmax + j] != exclude_id) ? contrib : 0.0;
                else e += contrib;
            }
        }
        return e * qq;
    }

    // ---------------------------------------------------------------- far field
    // kmc_octal.py:591-640
    double indirect(const double* p, double qq, int64_t lcell) const {
        const double* c = &centres[3 * lcell];
        double r, th, ph;
        spherical(p[0] - c[0], p[1] - c[1], p[2] - c[2], &r, &th, &ph);
        return qq * local_eval(B, &lexp[lcell * 2 * L * L], r, th, ph);
    }

    // ---------------------------------------------------------------- self
    // kmc_fmm_self_interaction.py:190-242
    double self_term(const double* op, double oq, const double* np_) const {
        const double coeff = oq * oq;
        const double dx = op[0] - np_[0], dy = op[1] - np_[1], dz = op[2] - np_[2];
        double e27 = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
        if (bc != BC_FREE) {
            e27 = self27;
            for (int k = 0; k < 27; k++) {
                const double ddx = E * OFF[k][0] + op[0] - np_[0];
                const double ddy = E * OFF[k][1] + op[1] - np_[1];
                const double ddz = E * OFF[k][2] + op[2] - np_[2];
                e27 += 1.0 / std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            }
        }
        if (bc == BC_FF) e27 = 0.0;
        return coeff * e27;
    }

    // ---------------------------------------------------------------- long range
    // kmc_full_long_range.py:516-560, 641-692
    double lr_apply_dot(const double* Mv, const double* Evv) const {
        const int L2 = L * L;
        double dot = 0.0;
        for (int row = 0; row < L2; row++) {
            double a = 0.0, b = 0.0;
            for (int64_t k = lr_indptr[row]; k < lr_indptr[row + 1]; k++) {
                a += lr_data[k] * Mv[lr_indices[k]];
                b += lr_data[k] * Mv[lr_indices[k] + L2];
            }
            dot += a * Evv[row] + b * Evv[row + L2];
        }
        const int re11 = 3, re10 = 2, re1n1 = 1;
        double d = 0.0;
        if (L > 1) {
            d += (dip[0] * Mv[re11]) * Evv[re11];
            d += (dip[0] * Mv[re11]) * Evv[re1n1];
            d -= (dip[1] * Mv[L2 + re11]) * Evv[L2 + re11];
            d += (dip[1] * Mv[L2 + re11]) * Evv[L2 + re1n1];
            d += (dip[2] * Mv[re10]) * Evv[re10];
        }
        return 0.5 * dot + 0.5 * d;
    }
    double lr_energy() const { return have_lr ? lr_apply_dot(M.data(), Ev.data()) : 0.0; }

    // per particle with its proposals : kmc_full_long_range.py:710-858
    void long_range(const double* op, double oq, int64_t nprop, const double* npos, const double* nq,
                    double old_energy, double* out) const {
        const int n2 = 2 * L * L;
        std::vector<double> om(M), oe(Ev), nm(n2), ne(n2);
        double r, th, ph;
        spherical(op[0], op[1], op[2], &r, &th, &ph);
        p2m(B, -1.0 * oq, r, th, ph, om.data());
        dot_vec(B, -1.0 * oq, r, th, ph, oe.data());
        for (int64_t t = 0; t < nprop; t++) {
            nm = om; ne = oe;
            spherical(npos[3 * t], npos[3 * t + 1], npos[3 * t + 2], &r, &th, &ph);
            p2m(B, nq[t], r, th, ph, nm.data());
            dot_vec(B, nq[t], r, th, ph, ne.data());
            out[t] += old_energy - lr_apply_dot(nm.data(), ne.data());
        }
    }

    // ---------------------------------------------------------------- propose
    // kmc_fmm.py:1001-1082 (+ diff_lib 583-663).  dE (eu-scaled) per proposal.
    int propose(int64_t P, const int64_t* pid, const int64_t* es, const double* npos,
                const double* nq_in, double* dE, double* comps) {
        const int64_t T = es[P];
        std::vector<double> nq(T);
        int err = 0;
        for (int64_t p = 0; p < P; p++)
            for (int64_t t = es[p]; t < es[p + 1]; t++) {
                nq[t] = nq_in ? nq_in[t] : q[pid[p]];
                // kmc_mpi_decomp.py:312-345 validity
                for (int d = 0; d < 3; d++) {
                    if (npos[3 * t + d] < -0.5 * E || npos[3 * t + d] > 0.5 * E) err++;
                    double dd = npos[3 * t + d] - pos[3 * pid[p] + d];
                    if (bc != BC_FREE) {
                        if (dd < -0.5 * E) dd += E;
                        if (dd > 0.5 * E) dd -= E;
                    }
                    if (std::fabs(dd) > max_move) err++;
                }
            }
        if (err) return err;
        std::vector<double> U1D(T, 0.0), U1I(T, 0.0), U0D(P, 0.0), U0I(P, 0.0), USI(T, 0.0);
        const bool nearfar = bc != BC_FF;
        if (nearfar) {
#pragma omp parallel for schedule(dynamic)
            for (int64_t t = 0; t < T; t++) {
                int64_t c3[3];
                bin(&npos[3 * t], c3);
                const int64_t lc = store_cell(lin(c3));
                U1D[t] = direct(&npos[3 * t], nq[t], lc, -1);
                U1I[t] = indirect(&npos[3 * t], nq[t], lc);
            }
#pragma omp parallel for schedule(dynamic)
            for (int64_t p = 0; p < P; p++) {
                const int64_t i = pid[p];
                const int64_t lc = store_cell(cell[i]);
                U0D[p] = direct(&pos[3 * i], q[i], lc, i);
                U0I[p] = indirect(&pos[3 * i], q[i], lc);
            }
        }
        const double old_lr = lr_energy();
#pragma omp parallel for schedule(dynamic)
        for (int64_t p = 0; p < P; p++) {
            const int64_t i = pid[p];
            for (int64_t t = es[p]; t < es[p + 1]; t++) USI[t] = self_term(&pos[3 * i], q[i], &npos[3 * t]);
            if (have_lr && (bc == BC_PBC || bc == BC_FF))
                long_range(&pos[3 * i], q[i], es[p + 1] - es[p], &npos[3 * es[p]], &nq[es[p]], old_lr, &USI[es[p]]);
        }
        for (int64_t p = 0; p < P; p++)
            for (int64_t t = es[p]; t < es[p + 1]; t++) {
                if (bc == BC_FF) dE[t] = eu * (-1.0 * USI[t]);
                else dE[t] = eu * ((U1D[t] + U1I[t] - U0D[p] - U0I[p]) - USI[t]);
                if (comps) {
                    comps[5 * t + 0] = U1D[t]; comps[5 * t + 1] = U1I[t];
                    comps[5 * t + 2] = U0D[p]; comps[5 * t + 3] = U0I[p];
                    comps[5 * t + 4] = USI[t];
                }
            }
        return 0;
    }

    // ---------------------------------------------------------------- checker
    // Proposal energies of the stored state (positions, cell store, expansions, M, E,
    // R, dipole terms: exactly the doubles the reference-ordered propose reads),
    // evaluated on the real type R.  With R = long double (64-bit mantissa) this is
    // the extended-precision reference the FP64 and mixed GPU paths are measured
    // against.  Near field, far field and self term follow direct / indirect /
    // self_term above.  The long-range term is not the reference's difference of two
    // whole-system energies (kmc_full_long_range.py:843-858, restated in long_range
    // above): with B the bilinear form of lr_apply_dot, M' = M + dM, E' = E + dE,
    //   E_lr(M, E) - E_lr(M', E') = -[ cE . dM + cM . dE + B(dM, dE) ],
    // cM = B(M, .) and cE = B(., E) as coefficient vectors, dM = P2M(new) - P2M(old)
    // and dE the same for the dot vector, so only quantities of the size of the moved
    // charge's own contribution are summed.
    template <typename R>
    R lr_form_t(const R* Mv, const R* Evv) const {
        const int L2 = L * L;
        R dot = 0;
        for (int row = 0; row < L2; row++) {
            R a = 0, b = 0;
            for (int64_t k = lr_indptr[row]; k < lr_indptr[row + 1]; k++) {
                a += R(lr_data[k]) * Mv[lr_indices[k]];
                b += R(lr_data[k]) * Mv[lr_indices[k] + L2];
            }
            dot += a * Evv[row] + b * Evv[row + L2];
        }
        R d = 0;
        if (L > 1) {
            const int re11 = 3, re10 = 2, re1n1 = 1;
            d += (R(dip[0]) * Mv[re11]) * Evv[re11];
            d += (R(dip[0]) * Mv[re11]) * Evv[re1n1];
            d -= (R(dip[1]) * Mv[L2 + re11]) * Evv[L2 + re11];
            d += (R(dip[1]) * Mv[L2 + re11]) * Evv[L2 + re1n1];
            d += (R(dip[2]) * Mv[re10]) * Evv[re10];
        }
        return R(0.5) * dot + R(0.5) * d;
    }
    // coefficient vectors cM[j] = B(M, e_j) and cE[k] = B(e_k, E) of the bilinear form
    template <typename R>
    void lr_linear_t(std::vector<R>& cM, std::vector<R>& cE) const {
        const int L2 = L * L, n2 = 2 * L2;
        std::vector<R> Mr(M.begin(), M.end()), Er(Ev.begin(), Ev.end()), e(n2, R(0));
        cM.assign(n2, R(0));
        cE.assign(n2, R(0));
        for (int j = 0; j < n2; j++) {
            e[j] = 1;
            cM[j] = lr_form_t<R>(Mr.data(), e.data());
            cE[j] = lr_form_t<R>(e.data(), Er.data());
            e[j] = 0;
        }
    }
    template <typename R>
    R direct_t(const double* p, R qq, int64_t lcell, int64_t exclude_id) const {
        R e = 0;
        for (int k = 0; k < 27; k++) {
            const int64_t jc = lcell + OFF[k][0] + (int64_t)lsd * (OFF[k][1] + (int64_t)lsd * OFF[k][2]);
            for (int64_t j = 0; j < occ[jc]; j++) {
                if (exclude_id >= 0 && pids[jc * occmax + j] == exclude_id) continue;
                const double* d = &pdata[(jc * occmax + j) * 5];
                const R dx = R(p[0]) - R(d[0]), dy = R(p[1]) - R(d[1]), dz = R(p[2]) - R(d[2]);
                e += R(d[3]) / std::sqrt(dx * dx + dy * dy + dz * dz);
            }
        }
        return e * qq;
    }
    template <typename R>
    R indirect_t(const BasisT<R>& Bt, const double* p, R qq, int64_t lcell) const {
        const double* c = &centres[3 * lcell];
        R r, th, ph;
        spherical_t<R>(R(p[0]) - R(c[0]), R(p[1]) - R(c[1]), R(p[2]) - R(c[2]), &r, &th, &ph);
        return qq * local_eval_t<R>(Bt, &lexp[lcell * 2 * L * L], r, th, ph);
    }
    template <typename R>
    R self_term_t(const double* op, R oq, const double* np_) const {
        if (bc == BC_FF) return 0;
        if (bc == BC_FREE) {
            const R dx = R(op[0]) - R(np_[0]), dy = R(op[1]) - R(np_[1]), dz = R(op[2]) - R(np_[2]);
            return oq * oq / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        R e27 = 0;
        for (int k = 0; k < 27; k++) {
            const R dx = R(E) * OFF[k][0] + R(op[0]) - R(np_[0]);
            const R dy = R(E) * OFF[k][1] + R(op[1]) - R(np_[1]);
            const R dz = R(E) * OFF[k][2] + R(op[2]) - R(np_[2]);
            e27 += R(1) / std::sqrt(dx * dx + dy * dy + dz * dz);
            if (OFF[k][0] != 0 || OFF[k][1] != 0 || OFF[k][2] != 0) {
                const R ex = R(E) * OFF[k][0], ey = R(E) * OFF[k][1], ez = R(E) * OFF[k][2];
                e27 -= R(1) / std::sqrt(ex * ex + ey * ey + ez * ez);
            }
        }
        return oq * oq * e27;
    }
    // dE (eu scaled) and optionally comps (T x 5: U1D, U1I, U0D, U0I, USI) on type R
    template <typename R>
    void propose_check(int64_t P, const int64_t* pid, const int64_t* es, const double* npos, const double* nq_in,
                       double* dE, double* comps) const {
        const BasisT<R> Bt(L);
        const bool nearfar = bc != BC_FF;
        const bool lr = have_lr && (bc == BC_PBC || bc == BC_FF);
        const int n2 = 2 * L * L;
        std::vector<R> cM, cE;
        if (lr) lr_linear_t<R>(cM, cE);
        std::vector<R> U0D(P, R(0)), U0I(P, R(0));
        std::vector<std::vector<R>> oM(lr ? P : 0), oE(lr ? P : 0);
#pragma omp parallel for schedule(dynamic)
        for (int64_t p = 0; p < P; p++) {
            const int64_t i = pid[p];
            if (nearfar) {
                const int64_t lc = store_cell(cell[i]);
                U0D[p] = direct_t<R>(&pos[3 * i], R(q[i]), lc, i);
                U0I[p] = indirect_t<R>(Bt, &pos[3 * i], R(q[i]), lc);
            }
            if (lr) {
                oM[p].assign(n2, R(0));
                oE[p].assign(n2, R(0));
                R r, th, ph;
                spherical_t<R>(R(pos[3 * i]), R(pos[3 * i + 1]), R(pos[3 * i + 2]), &r, &th, &ph);
                p2m_t<R>(Bt, R(q[i]), r, th, ph, oM[p].data(), false);
                p2m_t<R>(Bt, R(q[i]), r, th, ph, oE[p].data(), true);
            }
        }
#pragma omp parallel for schedule(dynamic)
        for (int64_t p = 0; p < P; p++) {
            const int64_t i = pid[p];
            std::vector<R> dM(n2), dEv(n2);
            for (int64_t t = es[p]; t < es[p + 1]; t++) {
                const R qn = nq_in ? R(nq_in[t]) : R(q[i]);
                R u1d = 0, u1i = 0;
                if (nearfar) {
                    int64_t c3[3];
                    bin(&npos[3 * t], c3);
                    const int64_t lc = store_cell(lin(c3));
                    u1d = direct_t<R>(&npos[3 * t], qn, lc, -1);
                    u1i = indirect_t<R>(Bt, &npos[3 * t], qn, lc);
                }
                R usi = self_term_t<R>(&pos[3 * i], R(q[i]), &npos[3 * t]);
                if (lr) {
                    for (int j = 0; j < n2; j++) { dM[j] = -oM[p][j]; dEv[j] = -oE[p][j]; }
                    R r, th, ph;
                    spherical_t<R>(R(npos[3 * t]), R(npos[3 * t + 1]), R(npos[3 * t + 2]), &r, &th, &ph);
                    p2m_t<R>(Bt, qn, r, th, ph, dM.data(), false);
                    p2m_t<R>(Bt, qn, r, th, ph, dEv.data(), true);
                    R lin_part = 0;
                    for (int j = 0; j < n2; j++) lin_part += cE[j] * dM[j] + cM[j] * dEv[j];
                    usi -= lin_part + lr_form_t<R>(dM.data(), dEv.data());
                }
                R de;
                if (bc == BC_FF) de = -usi;
                else de = ((u1d + u1i) - (U0D[p] + U0I[p])) - usi;
                dE[t] = (double)(R(eu) * de);
                if (comps) {
                    comps[5 * t + 0] = (double)u1d; comps[5 * t + 1] = (double)u1i;
                    comps[5 * t + 2] = (double)U0D[p]; comps[5 * t + 3] = (double)U0I[p];
                    comps[5 * t + 4] = (double)usi;
                }
            }
        }
    }

    // ---------------------------------------------------------------- accept
    // kmc_local.py:180-271 (extract), 274-359 (inject)
    int extract(int64_t gid, int64_t oldc) {
        int64_t c3[3];
        unlin(oldc, c3);
        std::vector<int> locs[3];
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < lsd; k++) {
                if (cell_ind[k] != c3[d]) continue;
                if (bc == BC_FREE && pfac[k] != 0) continue;
                if (bc == BC_27 && (pfac[k] > 1 || pfac[k] < -1)) continue;
                locs[d].push_back(k);
            }
        if (locs[0].empty() || locs[1].empty() || locs[2].empty()) return 0;
        const int64_t first = lidx(locs[0][0], locs[1][0], locs[2][0]);
        const int64_t oo = occ[first];
        int64_t index = -1, found = 0;
        for (int64_t s = 0; s < oo; s++)
            if (pids[first * occmax + s] == gid) { index = s; found++; }
        if (found != 1) return -1;   // "Index was not unique"
        for (int kx : locs[0]) for (int ky : locs[1]) for (int kz : locs[2]) {
            const int64_t li = lidx(kx, ky, kz);
            occ[li] = oo - 1;
            pids[li * occmax + index] = pids[li * occmax + oo - 1];
            for (int k = 0; k < 5; k++)
                pdata[(li * occmax + index) * 5 + k] = pdata[(li * occmax + oo - 1) * 5 + k];
        }
        return 0;
    }
    int inject(int64_t gid, int64_t newc, const double* np_, double qq) {
        int64_t c3[3];
        unlin(newc, c3);
        std::vector<int> locs[3];
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < lsd; k++) {
                if (cell_ind[k] != c3[d]) continue;
                if (bc == BC_FREE && pfac[k] != 0) continue;
                if (bc == BC_27 && (pfac[k] > 1 || pfac[k] < -1)) continue;
                locs[d].push_back(k);
            }
        if (locs[0].empty() || locs[1].empty() || locs[2].empty()) return 0;
        const int64_t first = lidx(locs[0][0], locs[1][0], locs[2][0]);
        const int64_t oo = occ[first];
        resize_store((int)oo + 1);
        for (int kx : locs[0]) for (int ky : locs[1]) for (int kz : locs[2]) {
            const int64_t li = lidx(kx, ky, kz);
            occ[li] = oo + 1;
            pids[li * occmax + oo] = gid;
            double* d = &pdata[(li * occmax + oo) * 5];
            d[0] = np_[0]; d[1] = np_[1]; d[2] = np_[2]; d[3] = qq;
            std::memcpy(&d[4], &gid, sizeof(int64_t));
            if (bc == BC_PBC || bc == BC_27) {
                d[0] += pfac[kx] * E; d[1] += pfac[ky] * E; d[2] += pfac[kz] * E;
            }
        }
        return 0;
    }

    // kmc_octal.py:1008-1089 : the hop case of accept_local_exp over every stored cell
    void accept_sweep(double qq, const double* np_, const double* op, int64_t newc, int64_t oldc) {
        int64_t nt[3], ot[3];
        unlin(newc, nt);
        unlin(oldc, ot);
        const int nimg = (bc == BC_FREE) ? 1 : 27;
        const int L2 = L * L;
#pragma omp parallel for schedule(dynamic) collapse(3)
        for (int kz = 0; kz < lsd; kz++)
            for (int ky = 0; ky < lsd; ky++)
                for (int kx = 0; kx < lsd; kx++) {
                    const int64_t li = lidx(kx, ky, kz);
                    std::vector<double> out(2 * L2, 0.0);
                    const double* c = &ocentres[3 * li];
                    const int64_t fc[3] = {cell_ind[kx], cell_ind[ky], cell_ind[kz]};
                    for (int pass = 0; pass < 2; pass++) {
                        const double* src = pass == 0 ? op : np_;
                        const int64_t* st = pass == 0 ? ot : nt;
                        const double qs = pass == 0 ? -1.0 * qq : qq;
                        for (int k = 0; k < nimg; k++) {
                            const int o[3] = {nimg == 1 ? 0 : OFF[k][0], nimg == 1 ? 0 : OFF[k][1], nimg == 1 ? 0 : OFF[k][2]};
                            bool near = true;
                            for (int d = 0; d < 3; d++) {
                                const int64_t dd = st[d] + o[d] * ncps - fc[d];
                                if (!(dd < 2 && dd > -2)) near = false;
                            }
                            if (near) continue;
                            double r, th, ph;
                            spherical(src[0] + o[0] * E - c[0], src[1] + o[1] * E - c[1], src[2] + o[2] * E - c[2], &r, &th, &ph);
                            p2l(B, qs, r, th, ph, out.data());
                        }
                    }
                    double* le = &lexp[li * 2 * L2];
                    for (int t = 0; t < 2 * L2; t++) le[t] += out[t];
                }
    }

    // P2L of every charge (and its 27 images for '27'/'pbc') into every stored cell:
    // the state the reference obtains from its initial FMM solve, without truncation
    // of the multipole side.
    void init_expansions() {
        const int nimg = (bc == BC_FREE) ? 1 : 27;
        const int L2 = L * L;
        std::fill(lexp.begin(), lexp.end(), 0.0);
        if (bc == BC_FF) return;
        std::vector<int64_t> c3s(3 * N);
        for (int64_t i = 0; i < N; i++) unlin(cell[i], &c3s[3 * i]);
#pragma omp parallel for schedule(dynamic) collapse(3)
        for (int kz = 0; kz < lsd; kz++)
            for (int ky = 0; ky < lsd; ky++)
                for (int kx = 0; kx < lsd; kx++) {
                    const int64_t li = lidx(kx, ky, kz);
                    double* le = &lexp[li * 2 * L2];
                    const double* c = &ocentres[3 * li];
                    const int64_t fc[3] = {cell_ind[kx], cell_ind[ky], cell_ind[kz]};
                    for (int64_t i = 0; i < N; i++)
                        for (int k = 0; k < nimg; k++) {
                            const int o[3] = {nimg == 1 ? 0 : OFF[k][0], nimg == 1 ? 0 : OFF[k][1], nimg == 1 ? 0 : OFF[k][2]};
                            bool near = true;
                            for (int d = 0; d < 3; d++) {
                                const int64_t dd = c3s[3 * i + d] + o[d] * ncps - fc[d];
                                if (!(dd < 2 && dd > -2)) near = false;
                            }
                            if (near) continue;
                            double r, th, ph;
                            spherical(pos[3 * i] + o[0] * E - c[0], pos[3 * i + 1] + o[1] * E - c[1],
                                      pos[3 * i + 2] + o[2] * E - c[2], &r, &th, &ph);
                            p2l(B, q[i], r, th, ph, le);
                        }
                }
    }

    // kmc_full_long_range.py:83-118
    void init_lr() {
        std::fill(M.begin(), M.end(), 0.0);
        std::fill(Ev.begin(), Ev.end(), 0.0);
        for (int64_t i = 0; i < N; i++) {
            double r, th, ph;
            spherical(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], &r, &th, &ph);
            p2m(B, q[i], r, th, ph, M.data());
            dot_vec(B, q[i], r, th, ph, Ev.data());
        }
    }

    // energy of the current state: 1/2 sum_i q_i phi(x_i) + long range
    double state_energy() const {
        double e = 0.0;
        if (bc != BC_FF) {
#pragma omp parallel for reduction(+ : e)
            for (int64_t i = 0; i < N; i++) {
                const int64_t lc = store_cell(cell[i]);
                e += 0.5 * (direct(&pos[3 * i], q[i], lc, i) + indirect(&pos[3 * i], q[i], lc));
            }
        }
        if (bc == BC_PBC || bc == BC_FF) e += lr_energy();
        return e;
    }

    // ---------------------------------------------------------------- inject / extract
    // kmc_inject_extract.py:82-124 (get_energy, per charge before the row sum):
    // old_energy_d (kmc_local.py:374-396, own id excluded), old_energy_i
    // (kmc_octal.py:363-375) and old_energy_l = q * local_eval(lrc(M), x)
    // (kmc_full_long_range.py:123-140, 901-960), whose bilinear form is the one
    // lr_apply_dot restates with the factor 1/2 removed.  eu applied.
    double charge_energy(int64_t id) const {
        const double* p = &pos[3 * id];
        double e = 0.0;
        if (bc != BC_FF) {
            const int64_t lc = store_cell(cell[id]);
            e += direct(p, q[id], lc, id) + indirect(p, q[id], lc);
        }
        if ((bc == BC_PBC || bc == BC_FF) && have_lr) e += lr_point(p, q[id]);
        return e * eu;
    }
    double lr_point(const double* p, double qq) const {
        std::vector<double> ex(2 * L * L, 0.0);
        double r, th, ph;
        spherical(p[0], p[1], p[2], &r, &th, &ph);
        dot_vec(B, qq, r, th, ph, ex.data());
        return 2.0 * lr_apply_dot(M.data(), ex.data());
    }
    // q * phi(x) at a point that holds no charge of the system (eval_field pieces,
    // kmc_fmm.py:544-567), eu applied
    double point_energy(const double* p, double qq) const {
        double e = 0.0;
        if (bc != BC_FF) {
            int64_t c3[3];
            bin(p, c3);
            const int64_t lc = store_cell(lin(c3));
            e += direct(p, qq, lc, -1) + indirect(p, qq, lc);
        }
        if ((bc == BC_PBC || bc == BC_FF) && have_lr) e += lr_point(p, qq);
        return e * eu;
    }
    // compute_energy (kmc_inject_extract.py:145-165) with the direct sums of
    // kmc_direct.py:150-300: nearest_direct (pairs once, plus half of every pair and
    // self pair over the 26 neighbour images for '27' / pbc) and pbc_direct (half the
    // linear operator and dipole correction applied to the group's own M and E)
    double group_energy(int64_t m, const double* gp, const double* gq) const {
        double e = 0.0;
        if (bc != BC_FF) {
            for (int64_t i = 0; i < m; i++) {
                for (int64_t j = i + 1; j < m; j++) {
                    const double dx = gp[3 * i] - gp[3 * j], dy = gp[3 * i + 1] - gp[3 * j + 1],
                                 dz = gp[3 * i + 2] - gp[3 * j + 2];
                    e += gq[i] * gq[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
                }
                if (bc == BC_27 || bc == BC_PBC)
                    for (int64_t j = 0; j < m; j++)
                        for (int k = 0; k < 27; k++) {
                            if (OFF[k][0] == 0 && OFF[k][1] == 0 && OFF[k][2] == 0) continue;
                            const double dx = gp[3 * j] - gp[3 * i] + OFF[k][0] * E;
                            const double dy = gp[3 * j + 1] - gp[3 * i + 1] + OFF[k][1] * E;
                            const double dz = gp[3 * j + 2] - gp[3 * i + 2] + OFF[k][2] * E;
                            e += 0.5 * gq[i] * gq[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
                        }
            }
        }
        if ((bc == BC_PBC || bc == BC_FF) && have_lr) {
            std::vector<double> gm(2 * L * L, 0.0), ge(2 * L * L, 0.0);
            for (int64_t i = 0; i < m; i++) {
                double r, th, ph;
                spherical(gp[3 * i], gp[3 * i + 1], gp[3 * i + 2], &r, &th, &ph);
                p2m(B, gq[i], r, th, ph, gm.data());
                dot_vec(B, gq[i], r, th, ph, ge.data());
            }
            e += lr_apply_dot(gm.data(), ge.data());
        }
        return e * eu;
    }

    // kmc_fmm.py:811-912
    int accept(int64_t gid, const double* np_, double* dE_out) {
        const double op[3] = {pos[3 * gid], pos[3 * gid + 1], pos[3 * gid + 2]};
        const double qq = q[gid];
        int64_t nc3[3];
        bin(np_, nc3);
        const int64_t newc = lin(nc3), oldc = cell[gid];
        const int64_t es[2] = {0, 1};
        double dE = 0.0;
        int err = propose(1, &gid, es, np_, nullptr, &dE, nullptr);
        if (err) return err;
        energy += dE;
        if (dE_out) *dE_out = dE;
        cell[gid] = newc;
        for (int d = 0; d < 3; d++) pos[3 * gid + d] = np_[d];
        if (bc != BC_FF) {
            if (extract(gid, oldc) < 0) return -1;
            inject(gid, newc, np_, qq);
            accept_sweep(qq, np_, op, newc, oldc);
        }
        if (bc == BC_PBC || bc == BC_FF) {
            double r, th, ph;
            spherical(op[0], op[1], op[2], &r, &th, &ph);
            p2m(B, -qq, r, th, ph, M.data());
            dot_vec(B, -qq, r, th, ph, Ev.data());
            spherical(np_[0], np_[1], np_[2], &r, &th, &ph);
            p2m(B, qq, r, th, ph, M.data());
            dot_vec(B, qq, r, th, ph, Ev.data());
        }
        return 0;
    }
};

}  // namespace

extern "C" {

void* orc_create(int64_t n, double extent, int32_t r, int32_t L, int32_t bc, double max_move, double eu) {
    static bool once = false;
    if (!once) { init_offsets(); once = true; }
    return new Oracle(n, extent, r, L, bc, max_move, eu);
}
void orc_destroy(void* h) { delete (Oracle*)h; }
void orc_set_threads(int32_t n) { if (n > 0) omp_set_num_threads(n); }
int32_t orc_get_threads() { return omp_get_max_threads(); }

void orc_set_particles(void* h, const double* pos, const double* q) { ((Oracle*)h)->set_particles(pos, q); }

// CSR of the reference-layout R (L^2 x L^2, applied to re and im halves) and (SX,SY,SZ)
void orc_set_lr(void* h, int64_t nnz, const double* data, const int64_t* indptr, const int64_t* indices,
                const double* dip3) {
    Oracle* o = (Oracle*)h;
    const int L2 = o->L * o->L;
    o->lr_data.assign(data, data + nnz);
    o->lr_indptr.assign(indptr, indptr + L2 + 1);
    o->lr_indices.assign(indices, indices + nnz);
    for (int d = 0; d < 3; d++) o->dip[d] = dip3[d];
    o->have_lr = true;
}

// P2L initial expansions (primary and halo copies), M/E and the state energy
void orc_initialise(void* h, double* energy) {
    Oracle* o = (Oracle*)h;
    o->init_expansions();
    o->init_lr();
    o->energy = o->state_energy() * o->eu;
    if (energy) *energy = o->energy;
}

// level-0 multipole M and dot vector E only (the expansions come from orc_set_expansions)
void orc_init_lr(void* h) { ((Oracle*)h)->init_lr(); }

// use given primary-cell expansions (reference layout [ncps^3][2L^2]) instead of P2L
void orc_set_expansions(void* h, const double* prim) {
    Oracle* o = (Oracle*)h;
    const int n2 = 2 * o->L * o->L;
    for (int kz = 0; kz < o->lsd; kz++)
        for (int ky = 0; ky < o->lsd; ky++)
            for (int kx = 0; kx < o->lsd; kx++) {
                const int64_t c3[3] = {o->cell_ind[kx], o->cell_ind[ky], o->cell_ind[kz]};
                std::memcpy(&o->lexp[o->lidx(kx, ky, kz) * n2], &prim[o->lin(c3) * n2], sizeof(double) * n2);
            }
}
void orc_get_expansions(void* h, double* prim) {
    Oracle* o = (Oracle*)h;
    const int n2 = 2 * o->L * o->L;
    for (int64_t c = 0; c < (int64_t)o->ncps * o->ncps * o->ncps; c++)
        std::memcpy(&prim[c * n2], &o->lexp[o->store_cell(c) * n2], sizeof(double) * n2);
}
// expansions of every stored (halo'd) cell, [lsd^3][2L^2]
void orc_get_store_expansions(void* h, double* all) {
    Oracle* o = (Oracle*)h;
    std::memcpy(all, o->lexp.data(), sizeof(double) * o->lexp.size());
}
void orc_get_ME(void* h, double* M, double* E) {
    Oracle* o = (Oracle*)h;
    std::memcpy(M, o->M.data(), sizeof(double) * o->M.size());
    std::memcpy(E, o->Ev.data(), sizeof(double) * o->Ev.size());
}
double orc_energy(void* h) { return ((Oracle*)h)->energy; }
void orc_set_energy(void* h, double e) { ((Oracle*)h)->energy = e; }
double orc_state_energy(void* h) { Oracle* o = (Oracle*)h; return o->state_energy() * o->eu; }
double orc_lr_energy(void* h) { return ((Oracle*)h)->lr_energy(); }
int32_t orc_lsd(void* h) { return ((Oracle*)h)->lsd; }
int32_t orc_occmax(void* h) { return ((Oracle*)h)->occmax; }

// primary-cell store as the reference keeps it: occ[ncps^3], ids[ncps^3][occmax], data[..][5]
void orc_get_store(void* h, int64_t* occ, int64_t* ids, double* data) {
    Oracle* o = (Oracle*)h;
    const int64_t nc = (int64_t)o->ncps * o->ncps * o->ncps;
    for (int64_t c = 0; c < nc; c++) {
        const int64_t li = o->store_cell(c);
        occ[c] = o->occ[li];
        for (int s = 0; s < o->occmax; s++) {
            ids[c * o->occmax + s] = o->pids[li * o->occmax + s];
            for (int k = 0; k < 5; k++) data[(c * o->occmax + s) * 5 + k] = o->pdata[(li * o->occmax + s) * 5 + k];
        }
    }
}
void orc_get_positions(void* h, double* pos) {
    Oracle* o = (Oracle*)h;
    std::memcpy(pos, o->pos.data(), sizeof(double) * o->pos.size());
}
void orc_get_cells(void* h, int64_t* cells) {
    Oracle* o = (Oracle*)h;
    std::memcpy(cells, o->cell.data(), sizeof(int64_t) * o->cell.size());
}

// tuple API: P moving charges pid[P], proposals es[p]..es[p+1] in npos (T x 3).
// dE (eu scaled) per proposal; comps (optional, T x 5) = U1D, U1I, U0D, U0I, USI.
int32_t orc_propose(void* h, int64_t P, const int64_t* pid, const int64_t* es, const double* npos,
                    const double* nq, double* dE, double* comps) {
    return ((Oracle*)h)->propose(P, pid, es, npos, nq, dE, comps);
}

// the checker: the proposal energies of the stored state on long double (extended != 0)
// or double, long-range term in the expanded difference form (Oracle::propose_check)
void orc_propose_check(void* h, int64_t P, const int64_t* pid, const int64_t* es, const double* npos,
                       const double* nq, double* dE, double* comps, int32_t extended) {
    const Oracle* o = (const Oracle*)h;
    if (extended) o->propose_check<long double>(P, pid, es, npos, nq, dE, comps);
    else o->propose_check<double>(P, pid, es, npos, nq, dE, comps);
}

// kmc_mpi_decomp.py:501-616 (setup_move) + kmc_fmm.py:665-762 (propose_with_dats)
int32_t orc_propose_with_dats(void* h, int64_t max_prop, const int64_t* site_max_counts,
                              const int64_t* current_sites, const double* prop_positions,
                              const int64_t* prop_masks, const double* prop_charges,
                              double* prop_energy_diffs) {
    Oracle* o = (Oracle*)h;
    std::vector<int64_t> es(1, 0), pid, rate;
    std::vector<double> npos, nq;
    for (int64_t px = 0; px < o->N; px++) {
        int64_t found = 0;
        const int64_t mx = site_max_counts[current_sites[px]];
        for (int64_t m = 0; m < max_prop && found < mx; m++)
            if (prop_masks[px * max_prop + m] > 0) {
                found++;
                for (int d = 0; d < 3; d++) npos.push_back(prop_positions[px * max_prop * 3 + m * 3 + d]);
                nq.push_back(prop_charges ? prop_charges[px * max_prop + m] : o->q[px]);
                rate.push_back(px * max_prop + m);
            }
        if (found > 0) { pid.push_back(px); es.push_back(es.back() + found); }
    }
    const int64_t P = (int64_t)pid.size();
    if (P == 0) return 0;
    std::vector<double> dE(es.back());
    const int err = o->propose(P, pid.data(), es.data(), npos.data(), nq.data(), dE.data(), nullptr);
    if (err) return err;
    for (size_t t = 0; t < dE.size(); t++) prop_energy_diffs[rate[t]] = dE[t];
    return 0;
}

int32_t orc_accept(void* h, int64_t pid, const double* new_pos, double* dE) {
    return ((Oracle*)h)->accept(pid, new_pos, dE);
}

// the accept sweep alone (for timing): applies the hop of charge q from op to np
void orc_accept_sweep_only(void* h, double q, const double* np_, const double* op) {
    Oracle* o = (Oracle*)h;
    int64_t n3[3], o3[3];
    o->bin(np_, n3);
    o->bin(op, o3);
    o->accept_sweep(q, np_, op, o->lin(n3), o->lin(o3));
}

// per-charge energies (get_energy before its row sum), eu applied
void orc_charge_energy(void* h, int64_t n, const int64_t* ids, double* out) {
    Oracle* o = (Oracle*)h;
    for (int64_t i = 0; i < n; i++) out[i] = o->charge_energy(ids[i]);
}
// q_i phi(x_i) at points without charges of the system, eu applied
void orc_point_energy(void* h, int64_t n, const double* p, const double* qq, double* out) {
    Oracle* o = (Oracle*)h;
    for (int64_t i = 0; i < n; i++) out[i] = o->point_energy(&p[3 * i], qq[i]);
}
// energy of m charges alone in the domain, eu applied
double orc_group_energy(void* h, int64_t m, const double* p, const double* qq) {
    return ((Oracle*)h)->group_energy(m, p, qq);
}

// basis primitives in the reference layout (for closed-form checks)
void orc_local_eval(int32_t L, const double* lexp, const double* disp, double* out) {
    Basis B(L);
    double r, th, ph;
    spherical(disp[0], disp[1], disp[2], &r, &th, &ph);
    *out = local_eval(B, lexp, r, th, ph);
}
void orc_p2l(int32_t L, double q, const double* disp, double* out) {
    Basis B(L);
    double r, th, ph;
    spherical(disp[0], disp[1], disp[2], &r, &th, &ph);
    p2l(B, q, r, th, ph, out);
}
void orc_p2m(int32_t L, double q, const double* disp, double* out) {
    Basis B(L);
    double r, th, ph;
    spherical(disp[0], disp[1], disp[2], &r, &th, &ph);
    p2m(B, q, r, th, ph, out);
}
void orc_dot_vec(int32_t L, double q, const double* disp, double* out) {
    Basis B(L);
    double r, th, ph;
    spherical(disp[0], disp[1], disp[2], &r, &th, &ph);
    dot_vec(B, q, r, th, ph, out);
}

}  // extern "C"
