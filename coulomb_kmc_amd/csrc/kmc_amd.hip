// MI355X (gfx950) FMM-KMC propose/accept.
//
// Device state of one system (all in HBM, owned by the context):
//   particles   pos4[n]  (x, y, z, q)          cell[n] (linear leaf index)
//   cell tables occ[ncells], slots[ncells*occmax] of (x, y, z, q) + id
//   expansions  lexp[ncells][L(L+1)/2] complex (m >= 0 only: for real charges the
//               m < 0 coefficients are the complex conjugates, so the reference's
//               [re L^2 | im L^2] block is reconstructed exactly on demand)
//   long range  M (packed like lexp), g = G M, E_lr = 0.5 M.g
//
// Kernels (the reference function each replaces):
//   k_near        direct_new / direct_old        coulomb_kmc/kmc_local.py:928-1031
//   k_far         indirect_interactions          coulomb_kmc/kmc_octal.py:495-707
//   k_combine     self_interaction + diff_lib    kmc_fmm_self_interaction.py:190-242,
//                                                kmc_fmm.py:583-663
//   k_lr_part,    long_range_energy              kmc_full_long_range.py:710-858
//   k_lr_prop2
//   k_sweep       accept_local_exp (hop)         coulomb_kmc/kmc_octal.py:1008-1089
//   k_accept_tail extract + inject,              coulomb_kmc/kmc_local.py:180-370,
//                 long range accept              kmc_full_long_range.py:270-316
//   k_dats_*      setup_move                     coulomb_kmc/kmc_mpi_decomp.py:501-616
//
// Solid harmonics are evaluated from Cartesian coordinates with the column
// recurrence documented in coulomb_kmc_amd/harmonics.py (no trigonometry); the
// coefficient tables are compile-time constants.

#include <mutex>

#include "kmc_device.h"

// ============================================================================
// kernels independent of the expansion order
// ============================================================================
__global__ void k_fill_i32(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// the largest leaf occupancy: wave maxima of a grid-stride pass, atomicMax into *out
// (which starts at 0)
__global__ void k_occ_peak(const int* __restrict__ occ, int ncells, int* __restrict__ out) {
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncells; i += gridDim.x * blockDim.x) m = max(m, occ[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

// ============================================================================
// near field by leaf tiles
// ============================================================================
// The evaluation points are counting-sorted by the 4x4x4 tile of leaves that holds
// them (k_tile_count, k_tile_scan, k_tile_scatter).  One workgroup per tile copies the
// charge records of the tile's 6x6x6 leaves (the tile and the leaves around it, the
// periodic images shifted by the box as the reference's halo copies are,
// coulomb_kmc/kmc_local.py:796-817) into LDS once, then evaluates every point of the
// tile from there (G lanes per point, the point's 27 leaves in stencil order, the
// slots of a leaf strided over the lanes, compensated sums): each record is read from
// memory once per tile instead of once per point whose stencil holds it.  The halo is
// staged leaf by leaf in x-fastest order, so a point's 27 leaves are 9 runs of 3
// consecutive leaves along x, each one contiguous range of the staged records, which
// the point's G lanes stride over.  A tile whose leaves hold more records than the LDS
// takes evaluates its points from memory.  Workgroups of 1024 threads: two fit on a
// CU (the LDS), 32 waves.
constexpr int NT_THREADS = 1024;
constexpr int NT_SLOTS = 2048;   // LDS records per tile (36 B each)
constexpr int NT_HALO = 216;     // 6^3 leaves

__device__ __forceinline__ int tile_of_cell(int c, int n) {
    const int nb = n >> 2;
    return ((c % n) >> 2) + nb * ((((c / n) % n) >> 2) + nb * ((c / (n * n)) >> 2));
}
// The per-tile counts are TSTRIDE ints apart, one per 128-byte line: the entries'
// atomics on counts packed into a few lines went to a few memory channels.
constexpr int TSTRIDE = 32;
// Each evaluation entry's rank among the entries of its tile (rank[x]) and the
// per-tile counts.  Entries come in moving-charge order, so a wave holds a few tiles:
// the lanes that share a tile take one atomic for all of them (the group's first lane
// adds the group's size), and each lane's rank is the returned base plus its place in
// the group.
__global__ void k_tile_count(int ne, const int32_t* __restrict__ cell, int n, int* __restrict__ tcount,
                             int* __restrict__ rank) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool on = x < ne;
    const int t = on ? tile_of_cell(cell[x], n) : -1;
    uint64_t todo = __ballot(on);
    int lead = lane;
    uint64_t grp = 0;
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;   // first lane not yet grouped
        const int tl = __shfl(t, l, 64);
        const uint64_t m = __ballot(t == tl) & todo;
        if (on && t == tl) { lead = l; grp = m; }
        todo &= ~m;
    }
    int base = 0;
    if (on && lane == lead) base = atomicAdd(&tcount[t * TSTRIDE], (int)__popcll(grp));
    base = __shfl(base, lead, 64);
    if (on) rank[x] = base + (int)__popcll(grp & ((1ull << lane) - 1ull));
}
// exclusive scan of the counts into toff[0..nt] (one workgroup: each thread a chunk of
// tiles, wave scans by shuffles, then the 16 wave totals); the counts back to zero
__global__ void __launch_bounds__(1024) k_tile_scan(int nt, int* __restrict__ tcount, int* __restrict__ toff) {
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int chunk = (nt + 1023) / 1024;
    const int b0 = min(nt, tid * chunk), b1 = min(nt, b0 + chunk);
    int s = 0;
    for (int g = b0; g < b1; g += 4) {   // (loads of a group of 4 issued together)
        int v4[4];
#pragma unroll
        for (int j = 0; j < 4; j++) v4[j] = g + j < b1 ? tcount[(g + j) * TSTRIDE] : 0;
#pragma unroll
        for (int j = 0; j < 4; j++) s += v4[j];
    }
    int v = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    if (wv == 0) {
        int w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    int c = v - s + (wv > 0 ? wsum[wv - 1] : 0);
    for (int i = b0; i < b1; i++) {
        const int k = tcount[i * TSTRIDE];
        toff[i] = c;
        c += k;
        tcount[i * TSTRIDE] = 0;
    }
    if (tid == 1023) toff[nt] = wsum[15];
}
__global__ void k_tile_scatter(int ne, const int32_t* __restrict__ cell, int n, const int* __restrict__ toff,
                               const int* __restrict__ rank, int* __restrict__ perm) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < ne) perm[toff[tile_of_cell(cell[x], n)] + rank[x]] = x;
}

// one point from memory (tiles whose leaves overflow the LDS): near_body's loop
template <int G, bool MIXED>
__device__ __forceinline__ CSum near_point_global(const EvalArgs& a, int e, int lane) {
    CSum acc;
    const double4 p = a.pos4[e];
    const int c = a.cell[e];
    const int ex = a.excl[e];
    const int n = a.g.ncps;
    const int tx = c % n, ty = (c / n) % n, tz = c / (n * n);
    const bool periodic = a.g.bc != KMC_BC_FREE_SPACE;
    for (int k = 0; k < 27; k++) {
        int ox, oy, oz;
        stencil(k, ox, oy, oz);
        int nx = tx + ox, ny = ty + oy, nz = tz + oz;
        double shx = 0.0, shy = 0.0, shz = 0.0;
        if (periodic) {
            if (nx < 0) { nx += n; shx = -a.g.E; } else if (nx >= n) { nx -= n; shx = a.g.E; }
            if (ny < 0) { ny += n; shy = -a.g.E; } else if (ny >= n) { ny -= n; shy = a.g.E; }
            if (nz < 0) { nz += n; shz = -a.g.E; } else if (nz >= n) { nz -= n; shz = a.g.E; }
        } else if (nx < 0 || nx >= n || ny < 0 || ny >= n || nz < 0 || nz >= n) {
            continue;
        }
        const int nc = nx + n * (ny + n * nz);
        const int occ = min(a.occ[nc], a.occmax);
        const size_t base = (size_t)nc * a.occmax;
        for (int s = lane; s < occ; s += G) {
            const double4 q = a.slot_pos[base + s];
            const double dx = p.x - (q.x + shx), dy = p.y - (q.y + shy), dz = p.z - (q.z + shz);
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (ex < 0 || a.slot_id[base + s] != ex) acc.add(q.w * rsqrt_sel<MIXED>(r2));
        }
    }
    return acc;
}

template <int G, bool MIXED>
__global__ void __launch_bounds__(NT_THREADS) k_near_tile(EvalArgs a, const int* __restrict__ toff,
                                                          const int* __restrict__ perm) {
    __shared__ double4 s_pos[NT_SLOTS];
    __shared__ int s_id[NT_SLOTS];
    __shared__ int s_start[NT_HALO + 1];
    __shared__ int s_leaf[NT_HALO];
    __shared__ double s_sh[NT_HALO][3];
    __shared__ int s_wsum[NT_THREADS / 64];
    const int tile = blockIdx.x;
    const int p0 = toff[tile], p1 = toff[tile + 1];
    if (p0 == p1) return;   // no point in this tile (uniform over the workgroup)
    const int n = a.g.ncps, nb = n >> 2;
    const int bx = tile % nb, by = (tile / nb) % nb, bz = tile / (nb * nb);
    const bool periodic = a.g.bc != KMC_BC_FREE_SPACE;
    const int tid = threadIdx.x, lane64 = tid & 63, wv = tid >> 6;
    // the 216 leaves: global leaf, image shift, occupancy; exclusive prefix of the
    // occupancies over the leaves (wave scans, then the wave totals)
    int occ_h = 0;
    if (tid < NT_HALO) {
        int g3[3] = {4 * bx + tid % 6 - 1, 4 * by + (tid / 6) % 6 - 1, 4 * bz + tid / 36 - 1};
        double sh[3] = {0.0, 0.0, 0.0};
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (g3[d] < 0) { if (periodic) { g3[d] += n; sh[d] = -a.g.E; } else ok = false; }
            else if (g3[d] >= n) { if (periodic) { g3[d] -= n; sh[d] = a.g.E; } else ok = false; }
        }
        const int leaf = ok ? g3[0] + n * (g3[1] + n * g3[2]) : 0;
        occ_h = ok ? min(a.occ[leaf], a.occmax) : 0;
        s_leaf[tid] = leaf;
        s_sh[tid][0] = sh[0];
        s_sh[tid][1] = sh[1];
        s_sh[tid][2] = sh[2];
    }
    int v = occ_h;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o, 64);
        if (lane64 >= o) v += y;
    }
    if (lane64 == 63) s_wsum[wv] = v;
    __syncthreads();
    int before = 0;
    for (int j = 0; j < wv; j++) before += s_wsum[j];
    if (tid < NT_HALO) s_start[tid] = before + v - occ_h;
    if (tid == NT_HALO - 1) s_start[NT_HALO] = before + v;
    __syncthreads();
    const int total = s_start[NT_HALO];
    const bool staged = total <= NT_SLOTS;
    if (staged) {
        // copy the records: four threads per halo leaf (864 of the 1024), each every
        // fourth slot of its leaf, two slots' loads issued at a time
        static_assert(4 * NT_HALO <= NT_THREADS, "four threads per halo leaf");
        if (tid < 4 * NT_HALO) {
            const int h = tid >> 2;
            const int d0 = s_start[h], cnt = s_start[h + 1] - d0;
            const size_t src = (size_t)s_leaf[h] * a.occmax;
            const double sx = s_sh[h][0], sy = s_sh[h][1], sz = s_sh[h][2];
            for (int j = tid & 3; j < cnt; j += 8) {
                const bool two = j + 4 < cnt;
                const double4 q0 = a.slot_pos[src + j];
                const int i0 = a.slot_id[src + j];
                double4 q1 = q0;
                int i1 = i0;
                if (two) {
                    q1 = a.slot_pos[src + j + 4];
                    i1 = a.slot_id[src + j + 4];
                }
                s_pos[d0 + j] = make_double4(q0.x + sx, q0.y + sy, q0.z + sz, q0.w);
                s_id[d0 + j] = i0;
                if (two) {
                    s_pos[d0 + j + 4] = make_double4(q1.x + sx, q1.y + sy, q1.z + sz, q1.w);
                    s_id[d0 + j + 4] = i1;
                }
            }
        }
        __syncthreads();
    }
    const int lane = tid % G;
    for (int q = tid / G; q < p1 - p0; q += NT_THREADS / G) {
        const int e = perm[p0 + q];
        CSum acc;
        if (staged) {
            const double4 p = a.pos4[e];
            const int c = a.cell[e];
            const int ex = a.excl[e];
            const int lx = c % n - 4 * bx, ly = (c / n) % n - 4 * by, lz = c / (n * n) - 4 * bz;
            // the 9 x-runs (leaves lx-1 .. lx+1 of halo row (ly+oy, lz+oz)): their ranges first
            int r0[9], r1[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const int h = lx + 6 * ((ly + k % 3) + 6 * (lz + k / 3));
                r0[k] = s_start[h];
                r1[k] = s_start[h + 3];
            }
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const int e1 = r1[k];
                for (int j = r0[k] + lane; j < e1; j += G) {
                    const double4 r = s_pos[j];
                    const double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
                    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    if (ex < 0 || s_id[j] != ex) acc.add(r.w * rsqrt_sel<MIXED>(r2));
                }
            }
        } else {
            acc = near_point_global<G, MIXED>(a, e, lane);
        }
        acc = csum_group<G>(acc);
        if (lane == 0) a.out[e] = acc.value() * a.pos4[e].w;
    }
}


template <int G, bool MIXED>
__global__ void __launch_bounds__(256) k_near(EvalArgs a) { near_body<G, MIXED>(a); }
template <int G, bool MIXED>
__global__ void __launch_bounds__(256) k_near_sorted(EvalArgs a, const int* __restrict__ perm) {
    near_sorted_body<G, MIXED>(a, perm);
}
// batched form: system blockIdx.z of an argument table (grid sized for the largest)
template <int G, bool MIXED>
__global__ void __launch_bounds__(256) kb_near(const EvalArgs* __restrict__ args) {
    near_body<G, MIXED>(args[blockIdx.z]);
}

template <bool MIXED>
__global__ void __launch_bounds__(256) k_near_cells(EvalArgs a) { near_cells_body<MIXED>(a); }
template <bool MIXED>
__global__ void __launch_bounds__(256) kb_near_cells(const EvalArgs* __restrict__ args) {
    near_cells_body<MIXED>(args[blockIdx.z]);
}

template <bool MIXED>
__global__ void __launch_bounds__(256) k_near_groups(EvalArgs a, int P, int T, const int64_t* __restrict__ es) {
    constexpr int KP = NEAR_GROUP_POINTS;
    constexpr int NB3 = NEAR_BOX * NEAR_BOX * NEAR_BOX;
    constexpr int SPL = 6;    // slots per lane per round
    __shared__ int s_start[4][NB3 + 1];
    __shared__ int s_cell[4][NB3];
    __shared__ int s_shift[4][NB3];
    __shared__ double s_pu[4][KP][3];    // point positions in the box frame
    __shared__ int4 s_pk[4][KP];         // first member leaf of the box per axis, excluded id
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + wv;
    if (p >= P) return;
    const int n = a.g.ncps;
    const double E = a.g.E;
    const bool periodic = a.g.bc != KMC_BC_FREE_SPACE;
    // es as k_setup validates it, clamped so that no entry index leaves [0, T + P)
    const int t0 = (int)min((int64_t)T, max((int64_t)0, es[p]));
    const int t1 = (int)min((int64_t)T, max((int64_t)t0, es[p + 1]));
    const int K = 1 + t1 - t0;
    const int c0 = a.cell[T + p];
    const int l0x = c0 % n, l0y = (c0 / n) % n, l0z = c0 / (n * n);
    for (int pass = 0; pass < K; pass += KP) {
        const int kn = min(KP, K - pass);
        // point of this lane (lanes < kn): position, leaf offset a from the old leaf,
        // unwrap count u, excluded id
        const int kk = pass + lane;
        const int e = lane < kn ? (kk == 0 ? T + p : t0 + kk - 1) : T + p;
        const double4 pt = a.pos4[e];
        const int ce = a.cell[e];
        const int ex = a.excl[e];
        int ax = ce % n - l0x, ay = (ce / n) % n - l0y, az = ce / (n * n) - l0z;
        if (periodic) {
            if (2 * ax > n) ax -= n; else if (2 * ax < -n) ax += n;
            if (2 * ay > n) ay -= n; else if (2 * ay < -n) ay += n;
            if (2 * az > n) az -= n; else if (2 * az < -n) az += n;
        }
        const int ux = floordiv(l0x + ax, n), uy = floordiv(l0y + ay, n), uz = floordiv(l0z + az, n);
        int mnx = lane < kn ? ax : 1 << 20, mny = lane < kn ? ay : 1 << 20, mnz = lane < kn ? az : 1 << 20;
        int mxx = lane < kn ? ax : -(1 << 20), mxy = lane < kn ? ay : -(1 << 20), mxz = lane < kn ? az : -(1 << 20);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mnx = min(mnx, __shfl_xor(mnx, o)); mny = min(mny, __shfl_xor(mny, o)); mnz = min(mnz, __shfl_xor(mnz, o));
            mxx = max(mxx, __shfl_xor(mxx, o)); mxy = max(mxy, __shfl_xor(mxy, o)); mxz = max(mxz, __shfl_xor(mxz, o));
        }
        const int bx = mxx - mnx + 3, by = mxy - mny + 3, bz = mxz - mnz + 3;
        if (bx <= NEAR_BOX && by <= NEAR_BOX && bz <= NEAR_BOX) {
            const int lox = mnx - 1, loy = mny - 1, loz = mnz - 1;
            const int ncb = bx * by * bz;
            // box leaves: slot counts, leaf ids, image shifts; exclusive scan of the counts
            int cntv[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int cb = lane + 64 * h;
                int cnt = 0;
                if (cb < ncb) {
                    const int ix = cb % bx, iy = (cb / bx) % by, iz = cb / (bx * by);
                    const int gx = l0x + lox + ix, gy = l0y + loy + iy, gz = l0z + loz + iz;
                    const int sx = floordiv(gx, n), sy = floordiv(gy, n), sz = floordiv(gz, n);
                    const bool inside = periodic || (sx == 0 && sy == 0 && sz == 0);
                    const int leaf = (gx - sx * n) + n * ((gy - sy * n) + n * (gz - sz * n));
                    cnt = inside ? min(a.occ[leaf], a.occmax) : 0;
                    s_cell[wv][cb] = inside ? leaf : 0;
                    s_shift[wv][cb] = (sx + 8) | ((sy + 8) << 4) | ((sz + 8) << 8);
                }
                cntv[h] = cnt;
            }
            int tot = 0;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                int v = cntv[h];
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int y = __shfl_up(v, o);
                    if (lane >= o) v += y;
                }
                const int cb = lane + 64 * h;
                if (cb < ncb) s_start[wv][cb] = tot + v - cntv[h];
                tot += __shfl(v, 63);
            }
            if (lane == 0) s_start[wv][ncb] = tot;
            if (lane < kn) {
                // box frame: a particle of box leaf b sits at its position + s_b E, the point
                // at its position + u E; the point's stencil is box leaves a-1 .. a+1
                s_pu[wv][lane][0] = pt.x + (double)ux * E;
                s_pu[wv][lane][1] = pt.y + (double)uy * E;
                s_pu[wv][lane][2] = pt.z + (double)uz * E;
                s_pk[wv][lane] = make_int4(ax - 1 - lox, ay - 1 - loy, az - 1 - loz, ex);
            }
            wave_sync();
            const int S = tot;
            double acc[KP];
#pragma unroll
            for (int k = 0; k < KP; k++) acc[k] = 0.0;
            for (int base = 0; base < S; base += 64 * SPL) {
                double qx[SPL], qy[SPL], qz[SPL], qw[SPL];
                int id[SPL], ixj[SPL], iyj[SPL], izj[SPL];
#pragma unroll
                for (int j = 0; j < SPL; j++) {
                    const int g0 = base + 64 * j + lane;
                    const bool ok = g0 < S;
                    const int g = ok ? g0 : S - 1;
                    int lo = 0, hi = ncb - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_start[wv][mid] <= g) lo = mid;
                        else hi = mid - 1;
                    }
                    const int cb = lo;
                    const size_t sidx = (size_t)s_cell[wv][cb] * a.occmax + (g - s_start[wv][cb]);
                    const double4 q = a.slot_pos[sidx];
                    const int sc = s_shift[wv][cb];
                    qx[j] = q.x + (double)((sc & 15) - 8) * E;
                    qy[j] = q.y + (double)(((sc >> 4) & 15) - 8) * E;
                    qz[j] = q.z + (double)(((sc >> 8) & 15) - 8) * E;
                    qw[j] = q.w;
                    id[j] = a.slot_id[sidx];
                    // slots past the list: a leaf index no stencil holds
                    ixj[j] = ok ? cb % bx : 64;
                    iyj[j] = (cb / bx) % by;
                    izj[j] = cb / (bx * by);
                }
#pragma unroll
                for (int k = 0; k < KP; k++) {
                    if (k < kn) {
                        const double pux = s_pu[wv][k][0], puy = s_pu[wv][k][1], puz = s_pu[wv][k][2];
                        const int4 pk = s_pk[wv][k];
#pragma unroll
                        for (int j = 0; j < SPL; j++) {
                            const bool member = (unsigned)(ixj[j] - pk.x) <= 2u && (unsigned)(iyj[j] - pk.y) <= 2u &&
                                                (unsigned)(izj[j] - pk.z) <= 2u;
                            const double dx = pux - qx[j], dy = puy - qy[j], dz = puz - qz[j];
                            double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                            const bool use = member && (pk.w < 0 || id[j] != pk.w);
                            r2 = use ? r2 : 1.0;
                            const double contrib = qw[j] * rsqrt_sel<MIXED>(r2);
                            acc[k] += use ? contrib : 0.0;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < KP; k++) {
                if (k < kn) {
                    double v = acc[k];
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                    if (lane == k) a.out[e] = v * pt.w;
                }
            }
            wave_sync();
        } else {
            // points spread over more leaves: each point over its own 27 leaves, lane k < 27
            // one leaf, its slots four at a time
            for (int k = 0; k < kn; k++) {
                const int ek = __builtin_amdgcn_readlane(e, k);
                const double px0 = readlane_d(pt.x, k), py0 = readlane_d(pt.y, k), pz0 = readlane_d(pt.z, k);
                const int ck = a.cell[ek];
                const int exk = __builtin_amdgcn_readlane(ex, k);
                double v = 0.0;
                if (lane < 27) {
                    const int tx = ck % n, ty = (ck / n) % n, tz = ck / (n * n);
                    int ox, oy, oz;
                    stencil(lane, ox, oy, oz);
                    int nx = tx + ox, ny = ty + oy, nz = tz + oz;
                    double shx = 0.0, shy = 0.0, shz = 0.0;
                    bool okc = true;
                    if (periodic) {
                        if (nx < 0) { nx += n; shx = -a.g.E; } else if (nx >= n) { nx -= n; shx = a.g.E; }
                        if (ny < 0) { ny += n; shy = -a.g.E; } else if (ny >= n) { ny -= n; shy = a.g.E; }
                        if (nz < 0) { nz += n; shz = -a.g.E; } else if (nz >= n) { nz -= n; shz = a.g.E; }
                    } else if (nx < 0 || nx >= n || ny < 0 || ny >= n || nz < 0 || nz >= n) {
                        okc = false;
                    }
                    if (okc) {
                        const int nc = nx + n * (ny + n * nz);
                        const int oc = min(a.occ[nc], a.occmax);
                        const size_t b0 = (size_t)nc * a.occmax;
                        for (int sl = 0; sl < oc; sl++) {
                            const double4 q = a.slot_pos[b0 + sl];
                            const double dx = (px0 - shx) - q.x, dy = (py0 - shy) - q.y, dz = (pz0 - shz) - q.z;
                            double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                            const bool use = exk < 0 || a.slot_id[b0 + sl] != exk;
                            r2 = use ? r2 : 1.0;
                            const double contrib = q.w * rsqrt_sel<MIXED>(r2);
                            v += use ? contrib : 0.0;
                        }
                    }
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                if (lane == k) a.out[e] = v * pt.w;
            }
        }
    }
}

__global__ void k_setup(SetupArgs a) { setup_body(a); }
// batched form: system blockIdx.z of an argument table
__global__ void kb_setup(const SetupArgs* __restrict__ args) { setup_body(args[blockIdx.z]); }

template <bool MIXED>
__global__ void __launch_bounds__(64) k_combine(CombineArgs a) { combine_body<MIXED>(a); }
template <bool MIXED>
__global__ void __launch_bounds__(64) kb_combine(const CombineArgs* __restrict__ args) {
    combine_body<MIXED>(args[blockIdx.z]);
}

// g = G M and E_lr = 0.5 M.g  (one block)
__global__ void __launch_bounds__(256) k_lr_refresh(int nrow, const double* gdat, const int64_t* gptr,
                                                     const int32_t* gidx, const double* M, double* g,
                                                     double* elr) {
    __shared__ double part[256];
    double acc = 0.0;
    for (int row = threadIdx.x; row < nrow; row += blockDim.x) {
        double y = 0.0;
        for (int64_t k = gptr[row]; k < gptr[row + 1]; k++) y = fma(gdat[k], M[gidx[k]], y);
        g[row] = y;
        acc = fma(M[row], y, acc);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *elr = 0.5 * part[0];
}

__global__ void k_reduce_rows(int nblk, int nrow, const double* partial, double* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partial[(size_t)b * nrow + r];
    out[r] = s;
}

__global__ void k_prepare_record(Geom g, int pid, double nx, double ny, double nz, const double4* ppos4,
                                 const int32_t* pcell, const double* dE, double* energy, MoveRec* rec) {
    const double4 o = ppos4[pid];
    rec->old_pos[0] = o.x; rec->old_pos[1] = o.y; rec->old_pos[2] = o.z;
    rec->new_pos[0] = nx; rec->new_pos[1] = ny; rec->new_pos[2] = nz;
    rec->q = o.w;
    rec->pid = pid;
    rec->old_cell = pcell[pid];
    const int n = g.ncps;
    rec->new_cell = bin_axis(nx, g.E, g.w, n) + n * (bin_axis(ny, g.E, g.w, n) + n * bin_axis(nz, g.E, g.w, n));
    rec->dE = dE[0];
    rec->valid = 1;
    *energy += dE[0];
}

__global__ void __launch_bounds__(1024) k_select(SelectArgs a) { select_body(a); }
__global__ void __launch_bounds__(1024) kb_select(const SelectArgs* __restrict__ args) {
    select_body(args[blockIdx.z]);
}

// a charge whose site index lies outside site_max_counts[nsites] takes no proposal and
// counts into err (the host returns KMC_E_ARG)
__global__ void k_dats_count(int n, int64_t max_prop, const int64_t* smc, int64_t nsites, const int64_t* sites,
                             const int64_t* masks, int64_t* cnt, int64_t* has, int* err) {
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= n) return;
    const int64_t site = sites[px];
    const bool ok = site >= 0 && site < nsites;
    if (!ok) atomicAdd(err, 1);
    const int64_t mx = ok ? smc[site] : 0;
    int64_t found = 0;
    for (int64_t m = 0; m < max_prop && found < mx; m++)
        if (masks[(int64_t)px * max_prop + m] > 0) found++;
    cnt[px] = found;
    has[px] = found > 0 ? 1 : 0;
}

// exclusive scan of n int64 values, three passes over blocks of SCAN_BLOCK values:
// per-block sums, a scan of the block sums (one workgroup, k_scan_excl), then each
// block's own scan plus its offset.  *total receives the sum.
constexpr int SCAN_BLOCK = 1024;
__global__ void __launch_bounds__(256) k_scan_block_sums(int n, const int64_t* __restrict__ in,
                                                         int64_t* __restrict__ bsum) {
    __shared__ int64_t part[4];
    const int b0 = blockIdx.x * SCAN_BLOCK, tid = threadIdx.x;
    int64_t v = 0;
    for (int i = b0 + tid; i < min(n, b0 + SCAN_BLOCK); i += 256) v += in[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((tid & 63) == 0) part[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) bsum[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}
__global__ void __launch_bounds__(256) k_scan_block_apply(int n, const int64_t* __restrict__ in,
                                                          const int64_t* __restrict__ boff, int64_t* __restrict__ out) {
    // each thread 4 consecutive values; wave scans, then the four wave totals
    __shared__ int64_t wsum[4];
    const int b0 = blockIdx.x * SCAN_BLOCK, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int64_t x[4], t = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = b0 + 4 * tid + j;
        x[j] = i < n ? in[i] : 0;
        t += x[j];
    }
    int64_t incl = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int64_t c = boff[blockIdx.x] + incl - t;
    for (int w = 0; w < wv; w++) c += wsum[w];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = b0 + 4 * tid + j;
        if (i < n) out[i] = c;
        c += x[j];
    }
}

// exclusive scan of n int64 values (single block, sequential chunks per thread)
__global__ void __launch_bounds__(1024) k_scan_excl(int n, const int64_t* in, int64_t* out, int64_t* total) {
    __shared__ int64_t sums[1024];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int chunk = (n + nt - 1) / nt;
    const int b0 = tid * chunk, b1 = min(n, b0 + chunk);
    int64_t s = 0;
    for (int i = b0; i < b1; i++) s += in[i];
    sums[tid] = s;
    __syncthreads();
    for (int o = 1; o < nt; o <<= 1) {
        const int64_t v = tid >= o ? sums[tid - o] : 0;
        __syncthreads();
        sums[tid] += v;
        __syncthreads();
    }
    int64_t c = tid > 0 ? sums[tid - 1] : 0;
    for (int i = b0; i < b1; i++) {
        out[i] = c;
        c += in[i];
    }
    if (tid == nt - 1) *total = sums[nt - 1];
}

__global__ void k_dats_scatter(int n, int64_t max_prop, const int64_t* smc, const int64_t* sites, const int64_t* masks,
                               const double* ppos, const double* pq, const int64_t* cnt, const int64_t* slot,
                               const int64_t* off, int64_t* pid, int64_t* es, double* new_pos, double* new_q,
                               int64_t* rate_loc, const int64_t* total_moves, const int64_t* total_parts,
                               const double4* ppos4) {
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px == 0) es[*total_parts] = *total_moves;
    if (px >= n || cnt[px] == 0) return;
    const int64_t p = slot[px];
    pid[p] = px;
    es[p] = off[px];
    const int64_t mx = smc[sites[px]];   // in range: k_dats_count gave cnt 0 otherwise
    int64_t found = 0;
    for (int64_t m = 0; m < max_prop && found < mx; m++)
        if (masks[(int64_t)px * max_prop + m] > 0) {
            const int64_t t = off[px] + found;
            for (int d = 0; d < 3; d++) new_pos[3 * t + d] = ppos[(int64_t)px * max_prop * 3 + m * 3 + d];
            new_q[t] = pq ? pq[(int64_t)px * max_prop + m] : ppos4[px].w;
            rate_loc[t] = (int64_t)px * max_prop + m;
            found++;
        }
}

__global__ void k_materialize(MatArgs a) { materialize_body(a); }
__global__ void kb_materialize(const MatArgs* __restrict__ args) { materialize_body(args[blockIdx.z]); }

__global__ void k_half_sum(int n, const double* a, const double* b, int nb, double* out) {
    // out = 0.5 * sum_i (a_i + sum_j b[j*n + i])  (one block, deterministic)
    __shared__ double part[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double bi = 0.0;
        for (int j = 0; j < nb; j++) bi += b[(size_t)j * n + i];
        s += a[i] + bi;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = 0.5 * part[0];
}

__global__ void k_finish_energy(const double* e_sr, const double* elr, double eu, int bc, double* energy) {
    double e = 0.0;
    if (bc != KMC_BC_FF_ONLY) e += *e_sr;
    if (bc == KMC_BC_PBC || bc == KMC_BC_FF_ONLY) e += *elr;
    *energy = e * eu;
}

// P2M straight from the particles of the leaves in each cell of a level:
// thread per (cell, packed (l, m)).
__global__ void __launch_bounds__(256) k_fmm_p2m(int L, FmmLevel lv, Geom g, const int* occ, const double4* slot_pos,
                                                 int occmax, const double* nrm, double2* Q) {
    const int nh = L * (L + 1) / 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ncell = (long long)lv.n * lv.n * lv.n;
    if (i >= ncell * nh) return;
    const int c = (int)(i / nh), h = (int)(i % nh);
    int l, m;
    half_lm(h, l, m);
    const int cx = c % lv.n, cy = (c / lv.n) % lv.n, cz = c / (lv.n * lv.n);
    const double x0 = -0.5 * g.E + (cx + 0.5) * lv.w, y0 = -0.5 * g.E + (cy + 0.5) * lv.w,
                 z0 = -0.5 * g.E + (cz + 0.5) * lv.w;
    const int f = g.ncps / lv.n;
    const double iw = 1.0 / lv.w;
    double sr = 0.0, si = 0.0;
    for (int lz = 0; lz < f; lz++)
        for (int ly = 0; ly < f; ly++)
            for (int lx = 0; lx < f; lx++) {
                const int leaf = (cx * f + lx) + g.ncps * ((cy * f + ly) + g.ncps * (cz * f + lz));
                const int o = occ[leaf];
                for (int j = 0; j < o; j++) {
                    const double4 p = slot_pos[(size_t)leaf * occmax + j];
                    const Cplx r = solid_lm<false>(l, m, (p.x - x0) * iw, (p.y - y0) * iw, (p.z - z0) * iw);
                    sr = fma(p.w, r.re, sr);
                    si = fma(-p.w, r.im, si);
                }
            }
    const double inv = 1.0 / nrm[hidx(l, m)];
    Q[(size_t)c * nh + h] = make_double2(sr * inv, si * inv);
}

// M2M: thread per (parent cell, packed (j, k)); rm[oct][n][mm + L - 1] = conj(R~_n^mm(b_oct))
__global__ void __launch_bounds__(256) k_fmm_m2m(int L, int np, const double2* Qc, const double2* rm, double2* Qp) {
    const int nh = L * (L + 1) / 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)np * np * np * nh) return;
    const int p = (int)(i / nh), h = (int)(i % nh);
    int j, k;
    half_lm(h, j, k);
    const int px = p % np, py = (p / np) % np, pz = p / (np * np);
    const int nc = 2 * np;
    double ar = 0.0, ai = 0.0;
    for (int oct = 0; oct < 8; oct++) {
        const int cx = 2 * px + (oct & 1), cy = 2 * py + ((oct >> 1) & 1), cz = 2 * pz + (oct >> 2);
        const double2* q = Qc + ((size_t)cx + (size_t)nc * (cy + (size_t)nc * cz)) * nh;
        const double2* t = rm + (size_t)oct * L * (2 * L - 1);
        for (int l = 0; l <= j; l++)
            for (int m = -l; m <= l; m++) {
                const int mm = k - m;
                if (mm > j - l || mm < -(j - l)) continue;
                const double2 a = half_get(q, l, m);
                const double2 b = t[(j - l) * (2 * L - 1) + mm + L - 1];
                ar = fma(a.x, b.x, fma(-a.y, b.y, ar));
                ai = fma(a.x, b.y, fma(a.y, b.x, ai));
            }
    }
    const double sc = ldexp(1.0, -j);
    Qp[(size_t)p * nh + h] = make_double2(ar * sc, ai * sc);
}

// L2L: thread per (child cell, packed (l, m)), l < Lout, from a parent expansion of
// order Lin >= Lout; rl[oct][n][mm + Lin - 1] = R~_n^mm(b_oct).  to_leaf: write the
// leaf-expansion layout L_lm = Lh_lm / (w^{l+1} N_lm) instead.
__global__ void __launch_bounds__(256) k_fmm_l2l(int Lin, int Lout, int nc, const double2* Lp, const double2* rl,
                                                 double2* Lc, int to_leaf, double w, const double* nrm) {
    const int nhi = Lin * (Lin + 1) / 2, nho = Lout * (Lout + 1) / 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nc * nc * nc * nho) return;
    const int c = (int)(i / nho), h = (int)(i % nho);
    int l, m;
    half_lm(h, l, m);
    const int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
    const int np = nc / 2;
    const int oct = (cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2);
    const double2* lp = Lp + ((size_t)(cx >> 1) + (size_t)np * ((cy >> 1) + (size_t)np * (cz >> 1))) * nhi;
    const int TW = 2 * Lin - 1;
    const double2* t = rl + (size_t)oct * Lin * TW;
    double ar = 0.0, ai = 0.0;
    for (int j = l; j < Lin; j++) {
        double sr = 0.0, si = 0.0;
        for (int k = -j; k <= j; k++) {
            const int mm = k - m;
            if (mm > j - l || mm < -(j - l)) continue;
            const double2 a = half_get(lp, j, k);
            const double2 b = t[(j - l) * TW + mm + Lin - 1];
            sr = fma(a.x, b.x, fma(-a.y, b.y, sr));
            si = fma(a.x, b.y, fma(a.y, b.x, si));
        }
        const double sc = ldexp(1.0, -(j + 1));
        ar = fma(sc, sr, ar);
        ai = fma(sc, si, ai);
    }
    if (to_leaf) {
        const double sc = 1.0 / (pow(w, (double)(l + 1)) * nrm[hidx(l, m)]);
        ar *= sc;
        ai *= sc;
    }
    Lc[(size_t)c * nho + h] = make_double2(ar, ai);
}

__global__ void __launch_bounds__(FMM_M2L_THREADS) k_fmm_m2l(int L, int n, int periodic, const double2* Q,
                                                             const double* nrm, double2* Lout) {
    extern __shared__ double2 fmm_lds[];
    const int nh = L * (L + 1) / 2;
    const int P = 2 * L - 1;             // orders 0 .. 2L-2
    const int QW = 4 * L - 3;            // q = -(2L-2) .. 2L-2
    double2* itab = fmm_lds;             // [P][QW]
    double2* qs = fmm_lds + P * QW;      // [nh]
    const int t = blockIdx.x;
    const int tx = t % n, ty = (t / n) % n, tz = t / (n * n);
    const int tid = threadIdx.x;
    int j = 0, k = 0;
    if (tid < nh) half_lm(tid, j, k);
    double ar = 0.0, ai = 0.0;
    for (int oz = -2 - (tz & 1); oz <= 3 - (tz & 1); oz++)
        for (int oy = -2 - (ty & 1); oy <= 3 - (ty & 1); oy++)
            for (int ox = -2 - (tx & 1); ox <= 3 - (tx & 1); ox++) {
                if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1 && oz >= -1 && oz <= 1) continue;
                int sx = tx + ox, sy = ty + oy, sz = tz + oz;
                if (periodic) {
                    sx = (sx + 2 * n) % n;
                    sy = (sy + 2 * n) % n;
                    sz = (sz + 2 * n) % n;
                } else if (sx < 0 || sx >= n || sy < 0 || sy >= n || sz < 0 || sz >= n) {
                    continue;
                }
                __syncthreads();   // previous source finished with the LDS tables
                if (tid < P) {
                    // column q = tid of I~_p^q(o), p = q .. 2L-2, by the irregular recurrence
                    const int q = tid;
                    const double x = ox, y = oy, z = oz;
                    const double r2 = x * x + y * y + z * z, ir2 = 1.0 / r2;
                    double sq = 1.0, pw = 1.0 / sqrt(r2);
                    Cplx wq = {1.0, 0.0};
                    for (int kk = 1; kk <= q; kk++) {
                        sq *= -sqrt((2.0 * kk - 1.0) / (2.0 * kk));
                        pw *= ir2;
                        wq = Cplx{wq.re * x - wq.im * y, wq.re * y + wq.im * x};
                    }
                    double b2 = 0.0, b1 = sq * pw;
                    for (int p = q; p < P; p++) {
                        if (p > q) {
                            const double b = (rec_alpha(p, q) * z * b1 - rec_beta(p, q) * b2) * ir2;
                            b2 = b1;
                            b1 = b;
                        }
                        const double f = b1 * nrm[hidx(p, q)];
                        const double2 e = make_double2(f * wq.re, f * wq.im);
                        itab[p * QW + q + P - 1] = e;
                        if (q > 0) itab[p * QW - q + P - 1] = conj_sgn(e, q);
                    }
                }
                const double2* src = Q + ((size_t)sx + (size_t)n * (sy + (size_t)n * sz)) * nh;
                for (int h = tid; h < nh; h += blockDim.x) qs[h] = src[h];
                __syncthreads();
                if (tid < nh) {
                    for (int l = 0; l < L; l++) {
                        const double sg = (l & 1) ? -1.0 : 1.0;
                        double sr = 0.0, si = 0.0;
                        for (int m = -l; m <= l; m++) {
                            const double2 a = itab[(j + l) * QW + (m - k) + P - 1];
                            const double2 b = m < 0 ? conj_sgn(qs[hidx(l, -m)], -m) : qs[hidx(l, m)];
                            sr = fma(a.x, b.x, fma(-a.y, b.y, sr));
                            si = fma(a.x, b.y, fma(a.y, b.x, si));
                        }
                        ar = fma(sg, sr, ar);
                        ai = fma(sg, si, ai);
                    }
                }
            }
    if (tid < nh) {
        const double sk = (k & 1) ? -1.0 : 1.0;
        double2 v = Lout[(size_t)t * nh + tid];
        v.x = fma(sk, ar, v.x);
        v.y = fma(sk, ai, v.y);
        Lout[(size_t)t * nh + tid] = v;
    }
}

// evaluation points: unit charge, leaf of the point, no exclusion
__global__ void k_eval_points(Geom g, int n, const double* pts, double4* pos4, int32_t* cell, int32_t* excl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    pos4[i] = make_double4(x, y, z, 1.0);
    const int nc = g.ncps;
    cell[i] = bin_axis(x, g.E, g.w, nc) + nc * (bin_axis(y, g.E, g.w, nc) + nc * bin_axis(z, g.E, g.w, nc));
    excl[i] = -1;
}

// long-range potential v(x) . g, v = packed multipole of a unit charge at x about
// the origin and g = G M: the first-order change of 0.5 M.(G M)
// (the reference evaluates its long-range local expansion at x,
// coulomb_kmc/kmc_full_long_range.py:341-387)
__global__ void k_eval_lr(int L, int n, const double4* pos4, const double* g, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos4[i];
    const double r2 = p.x * p.x + p.y * p.y + p.z * p.z;
    double acc = 0.0, s = 1.0;
    Cplx C = {1.0, 0.0};
    for (int m = 0; m < L; m++) {
        if (m > 0) {
            s *= -sqrt((2.0 * m - 1.0) / (2.0 * m));
            C = Cplx{C.re * p.x + C.im * p.y, C.im * p.x - C.re * p.y};   // (x - iy)^m
        }
        double b2 = 0.0, b1 = s;
        for (int l = m; l < L; l++) {
            if (l > m) {
                const double b = rec_alpha(l, m) * p.z * b1 - rec_beta(l, m) * r2 * b2;
                b2 = b1;
                b1 = b;
            }
            const int h = hidx(l, m);
            acc = fma(b1 * C.re, g[2 * h], fma(b1 * C.im, g[2 * h + 1], acc));
        }
    }
    out[i] = acc;
}

__global__ void k_eval_combine(int n, const double* ud, const double* ui, int nfar, const double* lr, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    if (ud) {
        v = ud[i];
        for (int j = 0; j < nfar; j++) v += ui[(size_t)j * n + i];
    }
    if (lr) v += lr[i];
    out[i] = v;
}

// charges by local id as evaluation entries (x, y, z, sign*q) in their leaf, with the id
// itself excluded from the near field (the old_* arrays of kmc_inject_extract.py:97-113).
// A bad id raises err bit 1 and leaves an inert zero-charge entry.
__global__ void k_gather_ids(Geom g, int n, const int64_t* ids, int64_t np, const double4* ppos4,
                             const int32_t* pcell, double sign, double4* pos4, int32_t* cell, int32_t* excl, int* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t id = ids[i];
    if (id < 0 || id >= np) {
        atomicOr(err, 1);
        pos4[i] = make_double4(g.c0, g.c0, g.c0, 0.0);
        cell[i] = 0;
        excl[i] = -1;
        return;
    }
    double4 p = ppos4[id];
    p.w *= sign;
    pos4[i] = p;
    cell[i] = pcell[id];
    excl[i] = (int32_t)id;
}

// near-field pair evaluations of the entries (the reference's direct-sum division /
// square-root count): the occupancies of the leaves around each entry's leaf
__global__ void k_pair_count(EvalArgs a, unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned int cnt = 0;
    if (i < a.n) {
        const int c = a.cell[i], n = a.g.ncps;
        const int tx = c % n, ty = (c / n) % n, tz = c / (n * n);
        for (int k = 0; k < 27; k++) {
            int ox, oy, oz;
            stencil(k, ox, oy, oz);
            int nx = tx + ox, ny = ty + oy, nz = tz + oz;
            if (a.g.bc != KMC_BC_FREE_SPACE) {
                nx = (nx + n) % n;
                ny = (ny + n) % n;
                nz = (nz + n) % n;
            } else if (nx < 0 || nx >= n || ny < 0 || ny >= n || nz < 0 || nz >= n) {
                continue;
            }
            cnt += (unsigned int)min(a.occ[nx + n * (ny + n * nz)], a.occmax);
        }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

// unit charge on every entry (the potential at the entries instead of their energy)
__global__ void k_unit_charges(int n, double4* pos4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pos4[i].w = 1.0;
}

// points with charges as evaluation entries, no exclusion
__global__ void k_pack_charges(Geom g, int n, const double* pos, const double* q, double4* pos4, int32_t* cell,
                               int32_t* excl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    pos4[i] = make_double4(x, y, z, q[i]);
    const int nc = g.ncps;
    cell[i] = bin_axis(x, g.E, g.w, nc) + nc * (bin_axis(y, g.E, g.w, nc) + nc * bin_axis(z, g.E, g.w, nc));
    if (excl) excl[i] = -1;
}

// eu * (near field + leaf expansion, both already carrying the charge, + q * long-range
// potential): old_energy_d + old_energy_i + old_energy_l of kmc_inject_extract.py:115-123
// (near/far/long-range parts: kmc_local.py:374-396, kmc_octal.py:363-375,
// kmc_full_long_range.py:123-140)
__global__ void k_charge_energy(int n, const double4* pos4, const double* ud, const double* ui, int nfar,
                                const double* lr, double eu, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    if (ud) {
        v = ud[i];
        for (int j = 0; j < nfar; j++) v += ui[(size_t)j * n + i];
    }
    if (lr) v = fma(pos4[i].w, lr[i], v);
    out[i] = eu * v;
}

// energy of each group of m charges alone in the domain (compute_energy,
// kmc_inject_extract.py:145-165; the direct sums it delegates to are restated from
// kmc_direct.py:150-300): every pair once; for '27' / pbc half of every pair and self
// pair over the 26 neighbour images; for pbc / ff-only 0.5 Mb.(G Mb) of the group's own
// level-0 multipole.  One workgroup per group; gdat == null: no long-range part.
__global__ void __launch_bounds__(256) k_group_energy(int m, const double4* __restrict__ pos4, double E, int bc,
                                                      int L, int nh, const double* gdat, const int64_t* gptr,
                                                      const int32_t* gidx, double eu, double* out) {
    extern __shared__ double Mb[];
    __shared__ double part[256];
    const double4* p = pos4 + (size_t)blockIdx.x * m;
    const int tid = threadIdx.x;
    double acc = 0.0;
    if (bc != 3) {
        const bool img = bc == 1 || bc == 2;
        const long long mm = (long long)m * m;
        for (long long t = tid; t < mm; t += blockDim.x) {
            const int i = (int)(t / m), j = (int)(t % m);
            const double4 a = p[i], b = p[j];
            const double qq = a.w * b.w;
            const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
            if (j > i) acc += qq / sqrt(dx * dx + dy * dy + dz * dz);
            if (img) {
                double s = 0.0;
                for (int k = 0; k < 27; k++) {
                    if (k == 13) continue;
                    const double ex = fma((double)(k % 3 - 1), E, dx);
                    const double ey = fma((double)((k / 3) % 3 - 1), E, dy);
                    const double ez = fma((double)(k / 9 - 1), E, dz);
                    s += 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
                }
                acc = fma(0.5 * qq, s, acc);
            }
        }
    }
    if (gdat) {
        for (int k = tid; k < 2 * nh; k += blockDim.x) Mb[k] = 0.0;
        __syncthreads();
        if (tid < L)
            for (int i = 0; i < m; i++) multipole_column_add(L, tid, p[i], Mb);
        __syncthreads();
        for (int row = tid; row < 2 * nh; row += blockDim.x) {
            double y = 0.0;
            for (int64_t k = gptr[row]; k < gptr[row + 1]; k++) y = fma(gdat[k], Mb[gidx[k]], y);
            acc = fma(0.5 * Mb[row], y, acc);
        }
    }
    part[tid] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) part[tid] += part[tid + o];
        __syncthreads();
    }
    if (tid == 0) out[blockIdx.x] = eu * part[0];
}

// propose_extract per row (kmc_inject_extract.py:178-213): energy of the row's charges
// alone (0 without self energy) minus the sum of their energies
__global__ void k_extract_rows(int nrows, int m, const double* ge, const double* ce, double* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double v = ge ? ge[r] : 0.0;
    for (int k = 0; k < m; k++) v -= ce[(size_t)r * m + k];
    out[r] = v;
}

// propose_inject (kmc_inject_extract.py:216-237): group energy + sum_i eu q_i phi(x_i)
__global__ void __launch_bounds__(256) k_inject_sum(int n, const double* ce, const double* ge, double* out) {
    __shared__ double part[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += ce[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = ge[0] + part[0];
}

__global__ void k_add_energy(const double* dE, double* energy) { *energy += *dE; }

// extraction from the cell tables and the particle arrays, one charge after another in
// descending id order: the cell's last occupant takes the freed slot (the store side of
// kmc_local.py:180-271), then the charge with the highest id takes the freed id (its
// slot entry renamed), so ids stay 0 .. n-1.  Flags bit 1 when an id is not found once.
__global__ void k_extract_store(int k, const int64_t* ids_desc, int n0, int* occ, double4* slot_pos, int32_t* slot_id,
                                int occmax, double4* ppos4, int32_t* pcell, int* flags) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int n = n0;
    for (int t = 0; t < k; t++) {
        const int id = (int)ids_desc[t];
        const int c = pcell[id];
        const size_t b = (size_t)c * occmax;
        const int oc = occ[c];
        int idx = -1, found = 0;
        for (int s = 0; s < oc; s++)
            if (slot_id[b + s] == id) { idx = s; found++; }
        if (found != 1) { atomicOr(flags, 1); return; }
        slot_id[b + idx] = slot_id[b + oc - 1];
        slot_pos[b + idx] = slot_pos[b + oc - 1];
        slot_id[b + oc - 1] = -1;
        occ[c] = oc - 1;
        const int last = n - 1;
        if (id != last) {
            const int lc = pcell[last];
            const size_t lb = (size_t)lc * occmax;
            const int lo = occ[lc];
            int li = -1, lf = 0;
            for (int s = 0; s < lo; s++)
                if (slot_id[lb + s] == last) { li = s; lf++; }
            if (lf != 1) { atomicOr(flags, 1); return; }
            slot_id[lb + li] = id;
            ppos4[id] = ppos4[last];
            pcell[id] = lc;
        }
        n = last;
    }
}

// injection: new charges take ids n0 .. n0+k-1 and are appended to their cells
// (the store side of kmc_local.py:274-359).  Flags bit 2 on a full cell.
__global__ void k_inject_store(int k, int n0, const double4* src, const int32_t* src_cell, int* occ, double4* slot_pos,
                               int32_t* slot_id, int occmax, double4* ppos4, int32_t* pcell, int* flags) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int t = 0; t < k; t++) {
        const int c = src_cell[t];
        const int o = occ[c];
        if (o >= occmax) { atomicOr(flags, 2); return; }
        const size_t sidx = (size_t)c * occmax + o;
        slot_id[sidx] = n0 + t;
        slot_pos[sidx] = src[t];
        occ[c] = o + 1;
        ppos4[n0 + t] = src[t];
        pcell[n0 + t] = c;
    }
}

// M += sum of per-block partial multipoles
__global__ void k_reduce_rows_add(int nblk, int nrow, const double* partial, double* M) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += partial[(size_t)b * nrow + r];
    M[r] += s;
}

__global__ void k_zero_i64(int64_t n, int64_t* p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// DiscoverInjectExtract (coulomb_kmc/kmc_inject_extract.py:439-533): each charge counts
// itself on every inject site within squared distance tol2 and records the first
// extract site within tol2 (1-based; 0: none)
__global__ void k_discover_sites(int64_t n, const double* pos, int ni, const double* isites, int ne,
                                 const double* esites, double tol2, unsigned long long* iflags, int64_t* eflag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    for (int s = 0; s < ni; s++) {
        const double dx = isites[3 * s] - px, dy = isites[3 * s + 1] - py, dz = isites[3 * s + 2] - pz;
        if (dx * dx + dy * dy + dz * dz < tol2) atomicAdd(&iflags[s], 1ull);
    }
    int64_t e = 0;
    for (int s = 0; s < ne; s++) {
        const double dx = esites[3 * s] - px, dy = esites[3 * s + 1] - py, dz = esites[3 * s + 2] - pz;
        if (dx * dx + dy * dy + dz * dz < tol2) { e = s + 1; break; }
    }
    eflag[i] = e;
}

// M proposals per charge: position + offset, wrapped into the domain as the example's
// prop_pos kernel does (below -E/2: +E, above E/2: -E), mask 1.  wrap == 0: no
// wrapping, proposals outside the domain get mask 0.
__global__ void k_lattice_proposals(int n, int M, double E, int wrap, const double4* ppos4, const double* offsets,
                                    double* new_pos, int64_t* masks) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * M) return;
    const int i = t / M, m = t - i * M;
    const double4 p = ppos4[i];
    const double lo = -0.5 * E, hi = 0.5 * E;
    double v[3] = {p.x + offsets[3 * m], p.y + offsets[3 * m + 1], p.z + offsets[3 * m + 2]};
    int64_t ok = 1;
    for (int d = 0; d < 3; d++) {
        if (wrap) {
            if (v[d] < lo) v[d] += E;
            if (v[d] > hi) v[d] -= E;
        } else if (v[d] < lo || v[d] > hi) {
            ok = 0;
        }
        new_pos[3 * (size_t)t + d] = v[d];
    }
    masks[t] = ok;
}

// overlap masking (the example's exclude PairLoop): proposal t of charge i gets mask 0
// when another charge lies within squared distance tol2 of it.  Periodic (wrap):
// minimum-image distance, as the PairLoop sees neighbours across the boundary.
// The charges are found in the 27 leaves around the proposal's leaf (tol2 <= w^2).
__global__ void k_overlap_mask(Geom g, int n, int M, double tol2, int wrap, const double* new_pos, const int* occ,
                               const double4* slot_pos, const int32_t* slot_id, int occmax, int64_t* masks) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * M) return;
    if (masks[t] <= 0) return;
    const int i = t / M;
    const double x = new_pos[3 * (size_t)t], y = new_pos[3 * (size_t)t + 1], z = new_pos[3 * (size_t)t + 2];
    const int nc = g.ncps;
    const int cx = bin_axis(x, g.E, g.w, nc), cy = bin_axis(y, g.E, g.w, nc), cz = bin_axis(z, g.E, g.w, nc);
    for (int k = 0; k < 27; k++) {
        int ox, oy, oz;
        stencil(k, ox, oy, oz);
        int jx = cx + ox, jy = cy + oy, jz = cz + oz;
        if (wrap) {
            jx = (jx + nc) % nc;
            jy = (jy + nc) % nc;
            jz = (jz + nc) % nc;
        } else if (jx < 0 || jy < 0 || jz < 0 || jx >= nc || jy >= nc || jz >= nc) {
            continue;
        }
        const int jc = jx + nc * (jy + nc * jz);
        const size_t b = (size_t)jc * occmax;
        const int o = occ[jc];
        for (int s = 0; s < o; s++) {
            if (slot_id[b + s] == i) continue;
            const double4 q = slot_pos[b + s];
            double dx = q.x - x, dy = q.y - y, dz = q.z - z;
            if (wrap) {
                dx -= g.E * rint(dx / g.E);
                dy -= g.E * rint(dy / g.E);
                dz -= g.E * rint(dz / g.E);
            }
            if (dx * dx + dy * dy + dz * dz < tol2) {
                masks[t] = 0;
                return;
            }
        }
    }
}

__global__ void k_lattice_index(int n, int M, int64_t* pid, int64_t* es) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pid[i] = i;
    if (i <= n) es[i] = (int64_t)i * M;
}

// one object per expansion order (kmc_order.hip, built with -DKMC_ORDER=L)
#define KMC_ORDERS(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) \
    X(18) X(19) X(20)
#define KMC_DECL(x) OrderFns kmc_order_fns_##x();
KMC_ORDERS(KMC_DECL)
#undef KMC_DECL
static bool order_fns(int L, OrderFns* f) {
    switch (L) {
#define KMC_CASE(x) case x: *f = kmc_order_fns_##x(); return true;
        KMC_ORDERS(KMC_CASE)
#undef KMC_CASE
        default: return false;
    }
}

template <bool MIXED>
static void launch_near_t(EvalArgs a, int G, hipStream_t s) {
    if (a.n <= 0) return;
    if (G == 27 || G == 0) {
        const long long threads = (long long)a.n * 32;
        hipLaunchKernelGGL((k_near_cells<MIXED>), dim3((int)((threads + 255) / 256)), dim3(256), 0, s, a);
        return;
    }
    const long long threads = (long long)a.n * G;
    const int blocks = (int)((threads + 255) / 256);
    switch (G) {
        case 1: hipLaunchKernelGGL((k_near<1, MIXED>), dim3(blocks), dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_near<4, MIXED>), dim3(blocks), dim3(256), 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_near<8, MIXED>), dim3(blocks), dim3(256), 0, s, a); break;
        case 16: hipLaunchKernelGGL((k_near<16, MIXED>), dim3(blocks), dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_near<32, MIXED>), dim3(blocks), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((k_near<64, MIXED>), dim3(blocks), dim3(256), 0, s, a); break;
    }
}
static void launch_near(EvalArgs a, int G, bool mixed, hipStream_t s) {
    if (G < 0) G = -G;   // the leaf-tiled layout needs the sorted points of a propose
    if (mixed) launch_near_t<true>(a, G, s);
    else launch_near_t<false>(a, G, s);
}
// evaluation entries of a propose (T proposals, then P old positions, grouped by es):
// the grouped kernel when selected (near_lanes 0), else the per-point kernels
static void launch_near_prop(EvalArgs a, int G, bool mixed, int P, int T, const int64_t* es, hipStream_t s) {
    if (G == 0 && es && P > 0) {
        const int blocks = (P + 3) / 4;
        if (mixed) hipLaunchKernelGGL(k_near_groups<true>, dim3(blocks), dim3(256), 0, s, a, P, T, es);
        else hipLaunchKernelGGL(k_near_groups<false>, dim3(blocks), dim3(256), 0, s, a, P, T, es);
        return;
    }
    launch_near(a, G, mixed, s);
}

// ============================================================================
// context
// ============================================================================
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t want) {
        if (want <= n && p) return hipSuccess;
        if (p) { hipFree(p); p = nullptr; n = 0; }
        if (want == 0) want = 1;
        hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
};

struct kmc_ctx {
    int64_t n = 0;
    int L = 0, r = 0, bc = 0, ncps = 0, nh = 0;
    int64_t ncells = 0;
    double E = 0, w = 0, max_move = 0, eu = 1.0, self27 = 0.0;
    int occ_headroom = 8;
    int occmax = 0;
    int near_lanes = 27;
    int near_order = 0;               // 1: per-point near field over the entries sorted by leaf tile
    int sweep_block = 0;
    int sweep_pairs = 2;
    int sweep_split = 1;
    int init_method = 0;              // 0 automatic, 1 direct P2L, 2 hierarchical
    int64_t last_P = 0, last_T = 0;   // sizes of the last propose (kernel timing)
    // side stream for work that overlaps within one call (long range beside the near and
    // far fields; cell tables and multipole beside the accept sweep), forked from and
    // joined back to the caller's stream with events (graph-capturable)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork[2] = {nullptr, nullptr}, ev_join[2] = {nullptr, nullptr};
    // second side stream for the accept sweep of a pipelined run of steps: the sweep of
    // step i runs beside the set-up, near field and long range of step i + 1
    hipStream_t side2 = nullptr;
    hipEvent_t ev_fork2 = nullptr, ev_sweep = nullptr;
    hipEvent_t ev_setup = nullptr, ev_far = nullptr;   // pipelined mode 3: far field on side2
    bool have_particles = false, have_lr = false, initialised = false;
    // the host mirrors below lag the device after device-side steps
    bool mirrors_stale = false;
    // bumped whenever device buffers are reallocated or the charge count changes
    // (batches recorded against the old layout refuse to step)
    int64_t generation = 0;
    OrderFns fns{};
    // kmc_profile_*: device time of the hot kernels of host-called (not captured) proposes
    // and accepts, HIP event pairs summed when read
    struct ProfPair {
        hipEvent_t a, b;
        int kind;
        int64_t count;
    };
    bool prof = false;
    std::vector<ProfPair> prof_pending;
    std::vector<hipEvent_t> prof_free;
    double prof_ms[4] = {0, 0, 0, 0};       // near field, far field, long range, accept sweep
    int64_t prof_count[4] = {0, 0, 0, 0};   // evaluation points (near, far), proposals (long range), sweeps
    DBuf<unsigned long long> prof_pairs;    // near-field pair evaluations (one rsqrt each)

    // particles
    DBuf<double4> ppos4;
    DBuf<int32_t> pcell;
    std::vector<double> h_pos, h_q;
    std::vector<int32_t> h_cell;
    std::vector<int32_t> h_occ;
    // cell tables
    DBuf<int> occ;
    DBuf<double4> slot_pos;
    DBuf<int32_t> slot_id;
    // expansions
    DBuf<double2> lexp;
    // long range
    DBuf<double> M, gvec, gdat, elr, gblk, aold, cold;
    DBuf<int64_t> gptr;
    DBuf<int32_t> gidx;
    int64_t gnnz = 0;
    // scalars
    DBuf<double> energy;
    DBuf<int> flags, err;
    DBuf<int> occpeak;                // kmc_store_reserve's largest occupancy
    DBuf<MoveRec> rec;
    // propose work buffers
    DBuf<double4> epos4;
    DBuf<int32_t> ecell, eexcl, owner;
    DBuf<double> UD, UI, lrd, dE1, rates, bsum;
    DBuf<int64_t> one_pid, one_es;
    DBuf<int64_t> es_copy;            // es of the last propose set-up
    DBuf<int> tcount, toff, rank, perm;   // leaf-tiled near field (counting sort by tile)
    DBuf<double> one_pos;
    // propose_with_dats work buffers
    DBuf<int64_t> d_cnt, d_has, d_slot, d_off, d_pid, d_es, d_rate, d_tot, d_scan;
    DBuf<double> d_npos, d_nq;
    // lattice steps: every charge proposes the same M offsets
    DBuf<int64_t> lat_pid, lat_es;
    DBuf<double> lat_off;
    std::vector<double> h_lat_off;
    int64_t lat_n = -1;

    Geom geom() const { return Geom{E, w, -0.5 * E + 0.5 * w, ncps, bc}; }
};

// leaf tiles (4x4x4) of the context's grid, 0 when the grid is not a multiple of 4 per side
static int near_tiles(const kmc_ctx* c) {
    return (c->ncps % 4 == 0) ? (c->ncps / 4) * (c->ncps / 4) * (c->ncps / 4) : 0;
}

static int check_ctx(kmc_ctx* c) {
    if (!c) return set_err(KMC_E_ARG, "null context");
    return KMC_OK;
}

// An error found through the device error counter: the counter goes back to zero (the
// device-side steps read it in k_select), then the error is returned.
static int fail_err_counter(kmc_ctx* c, hipStream_t s, int code, const char* msg) {
    HIPCHK(hipMemsetAsync(c->err.p, 0, sizeof(int), s));
    HIPCHK(hipStreamSynchronize(s));
    return set_err(code, msg);
}

static int ensure_eval(kmc_ctx* c, int64_t T, int64_t P) {
    const size_t ne = (size_t)(T + P);
    HIPCHK(c->epos4.ensure(ne));
    HIPCHK(c->ecell.ensure(ne));
    HIPCHK(c->eexcl.ensure(ne));
    HIPCHK(c->owner.ensure((size_t)std::max<int64_t>(T, 1)));
    HIPCHK(c->UD.ensure(ne));
    HIPCHK(c->UI.ensure(ne * (size_t)far_parts(c->L)));
    HIPCHK(c->lrd.ensure((size_t)12 * std::max<int64_t>(T, 1)));
    HIPCHK(c->rates.ensure((size_t)std::max<int64_t>(T, 1)));
    HIPCHK(c->bsum.ensure((size_t)std::max<int64_t>((T + 63) / 64, 1)));
    HIPCHK(c->es_copy.ensure((size_t)P + 1));
    const int nt = near_tiles(c);
    if (nt > 0) {
        const bool fresh = c->tcount.n < (size_t)nt * TSTRIDE;
        HIPCHK(c->tcount.ensure((size_t)nt * TSTRIDE));
        if (fresh) HIPCHK(hipMemset(c->tcount.p, 0, sizeof(int) * nt * TSTRIDE));   // the scan keeps it zero after
        HIPCHK(c->toff.ensure((size_t)nt + 1));
        HIPCHK(c->rank.ensure(ne));
        HIPCHK(c->perm.ensure(ne));
    }
    return KMC_OK;
}

// The evaluation arrays of one propose (T new points, then P old): the context's own
// buffers (ctx_work, sized by ensure_eval / ensure_lr_work) or a carve of the library's
// fixed scratch array (kmc_propose_noalloc).
struct ProposeWork {
    double4* epos4;
    int32_t *ecell, *eexcl, *owner;
    double *UD, *UI, *lrd, *aold, *cold;
    int64_t* es_copy;
    int* err;
    // leaf-tiled near field: per-tile counts (zero between calls), offsets, each
    // entry's rank in its tile, and the evaluation entries sorted by tile
    int *tcount, *toff, *rank, *perm;
};
static ProposeWork ctx_work(kmc_ctx* c) {
    return ProposeWork{c->epos4.p, c->ecell.p, c->eexcl.p, c->owner.p, c->UD.p,   c->UI.p,   c->lrd.p,
                       c->aold.p,  c->cold.p,  c->es_copy.p, c->err.p, c->tcount.p, c->toff.p, c->rank.p,
                       c->perm.p};
}

// near field of a propose's evaluation entries: the leaf-tiled kernel (near_lanes < 0,
// grids with a multiple of 4 leaves per side), the grouped one (0) or the per-point ones
template <bool MIXED>
static void launch_near_tiled(const kmc_ctx* c, hipStream_t s, const EvalArgs& a, int G, const ProposeWork& w) {
    const int n = c->ncps, nt = near_tiles(c), ne = a.n;
    hipLaunchKernelGGL(k_tile_count, dim3((ne + 255) / 256), dim3(256), 0, s, ne, a.cell, n, w.tcount, w.rank);
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, nt, w.tcount, w.toff);
    hipLaunchKernelGGL(k_tile_scatter, dim3((ne + 255) / 256), dim3(256), 0, s, ne, a.cell, n, w.toff, w.rank, w.perm);
    switch (G) {
        case 4: hipLaunchKernelGGL((k_near_tile<4, MIXED>), dim3(nt), dim3(NT_THREADS), 0, s, a, w.toff, w.perm); break;
        case 8: hipLaunchKernelGGL((k_near_tile<8, MIXED>), dim3(nt), dim3(NT_THREADS), 0, s, a, w.toff, w.perm); break;
        default: hipLaunchKernelGGL((k_near_tile<16, MIXED>), dim3(nt), dim3(NT_THREADS), 0, s, a, w.toff, w.perm); break;
    }
}
// the counting sort of the entries by leaf tile, then the per-point kernel over that order
template <bool MIXED>
static void launch_near_sorted(const kmc_ctx* c, hipStream_t s, const EvalArgs& a, int G, const ProposeWork& w) {
    const int n = c->ncps, nt = near_tiles(c), ne = a.n;
    hipLaunchKernelGGL(k_tile_count, dim3((ne + 255) / 256), dim3(256), 0, s, ne, a.cell, n, w.tcount, w.rank);
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, nt, w.tcount, w.toff);
    hipLaunchKernelGGL(k_tile_scatter, dim3((ne + 255) / 256), dim3(256), 0, s, ne, a.cell, n, w.toff, w.rank, w.perm);
    const dim3 grid((unsigned)(((long long)ne * G + 255) / 256));
    switch (G) {
        case 1: hipLaunchKernelGGL((k_near_sorted<1, MIXED>), grid, dim3(256), 0, s, a, w.perm); break;
        case 4: hipLaunchKernelGGL((k_near_sorted<4, MIXED>), grid, dim3(256), 0, s, a, w.perm); break;
        case 8: hipLaunchKernelGGL((k_near_sorted<8, MIXED>), grid, dim3(256), 0, s, a, w.perm); break;
        case 16: hipLaunchKernelGGL((k_near_sorted<16, MIXED>), grid, dim3(256), 0, s, a, w.perm); break;
        case 32: hipLaunchKernelGGL((k_near_sorted<32, MIXED>), grid, dim3(256), 0, s, a, w.perm); break;
        default: hipLaunchKernelGGL((k_near_sorted<64, MIXED>), grid, dim3(256), 0, s, a, w.perm); break;
    }
}
static void launch_near_any(const kmc_ctx* c, hipStream_t s, const EvalArgs& a, bool mixed, int P, int T,
                            const ProposeWork& w) {
    if (a.n <= 0) return;
    const int G = c->near_lanes;
    if (c->near_order == 1 && (G == 1 || G == 4 || G == 8 || G == 16 || G == 32 || G == 64) && near_tiles(c) > 0 &&
        w.perm) {
        if (mixed) launch_near_sorted<true>(c, s, a, G, w);
        else launch_near_sorted<false>(c, s, a, G, w);
        return;
    }
    if (c->near_lanes < 0 && near_tiles(c) > 0 && w.perm) {
        if (mixed) launch_near_tiled<true>(c, s, a, -c->near_lanes, w);
        else launch_near_tiled<false>(c, s, a, -c->near_lanes, w);
        return;
    }
    launch_near_prop(a, c->near_lanes, mixed, P, T, w.es_copy, s);
}

static EvalArgs eval_args(kmc_ctx* c, int n, const double4* pos4, const int32_t* cell, const int32_t* excl, double* out) {
    EvalArgs a;
    a.g = c->geom();
    a.n = n;
    a.pos4 = pos4;
    a.cell = cell;
    a.excl = excl;
    a.out = out;
    a.lexp = c->lexp.p;
    a.occ = c->occ.p;
    a.slot_pos = c->slot_pos.p;
    a.slot_id = c->slot_id.p;
    a.occmax = c->occmax;
    a.ppos4 = c->ppos4.p;
    a.np = (int)c->n;
    return a;
}

// row chunks of each operator block in k_lr_part: with few moving charges (under 4096)
// one lane per (charge, block) leaves most of the device idle
static int lr_split(int64_t P, const ProposeWork* = nullptr) { return P < 4096 ? LR_SPLIT_MAX : 1; }

static int ensure_lr_work(kmc_ctx* c, int64_t P) {
    HIPCHK(c->aold.ensure((size_t)12 * P * lr_nbmax(c->L)));
    HIPCHK(c->cold.ensure((size_t)12 * P * lr_split(P, nullptr)));
    return KMC_OK;
}

static LR2Args lr_args(kmc_ctx* c, int64_t P, int64_t T, const ProposeWork* w = nullptr) {
    const ProposeWork wk = w ? *w : ctx_work(c);
    LR2Args la;
    la.P = (int)P;
    la.T = (int)T;
    la.gblk = c->gblk.p;
    la.gvec = c->gvec.p;
    la.epos4 = wk.epos4;
    la.owner = wk.owner;
    la.aold = wk.aold;
    la.cpart = wk.cold;
    la.rsplit = lr_split(P, w);
    la.lpart = wk.lrd;
    return la;
}

// propose kernels after the evaluation arrays are filled (T new points, then P old)
// ---------------------------------------------------------------------------
// kernel arguments of one context (shared by the single-system and batched paths)
// ---------------------------------------------------------------------------
static bool uses_lr(const kmc_ctx* c) { return (c->bc == KMC_BC_PBC || c->bc == KMC_BC_FF_ONLY) && c->have_lr; }

static SetupArgs setup_args(kmc_ctx* c, int64_t P, int64_t T, const int64_t* d_pid, const int64_t* d_es,
                            const double* d_new_pos, const double* d_new_q, const ProposeWork* w = nullptr) {
    const ProposeWork wk = w ? *w : ctx_work(c);
    SetupArgs a;
    a.g = c->geom();
    a.max_move = c->max_move;
    a.P = (int)P;
    a.T = (int)T;
    a.n = (int)c->n;
    a.pid = d_pid;
    a.es = d_es;
    a.new_pos = d_new_pos;
    a.new_q = d_new_q;
    a.ppos4 = c->ppos4.p;
    a.pcell = c->pcell.p;
    a.epos4 = wk.epos4;
    a.ecell = wk.ecell;
    a.eexcl = wk.eexcl;
    a.owner = wk.owner;
    a.err = wk.err;
    a.es_out = wk.es_copy;
    a.mask = nullptr;
    return a;
}

static CombineArgs combine_args(kmc_ctx* c, int64_t P, int64_t T, double* d_out, const int64_t* rate_loc, double kT,
                                double* rates, const ProposeWork* w = nullptr) {
    const ProposeWork wk = w ? *w : ctx_work(c);
    const bool lr = uses_lr(c);
    CombineArgs ca;
    ca.g = c->geom();
    ca.T = (int)T;
    ca.eu = c->eu;
    ca.self27 = c->self27;
    ca.epos4 = wk.epos4;
    ca.owner = wk.owner;
    ca.UD = wk.UD;
    ca.UI = wk.UI;
    ca.nfar = far_parts(c->L);
    ca.lpart = lr ? wk.lrd : nullptr;
    ca.cpart = lr ? wk.cold : nullptr;
    ca.ncpart = 12 * lr_split(P, w);
    ca.P = (int)P;
    ca.rate_loc = rate_loc;
    ca.out = d_out;
    ca.kT = kT;
    ca.rates = rates;
    ca.bsum = rates ? c->bsum.p : nullptr;
    ca.mask = nullptr;
    ca.rate_rule = 0;
    return ca;
}

static TailArgs tail_args(kmc_ctx* c, int64_t* d_counter) {
    TailArgs ta;
    ta.rec = c->rec.p;
    ta.occ = c->occ.p;
    ta.slot_pos = c->slot_pos.p;
    ta.slot_id = c->slot_id.p;
    ta.occmax = c->occmax;
    ta.ppos4 = c->ppos4.p;
    ta.pcell = c->pcell.p;
    ta.flags = c->flags.p;
    ta.lr = uses_lr(c) ? 1 : 0;
    ta.M = c->M.p;
    ta.gdat = c->gdat.p;
    ta.gptr = c->gptr.p;
    ta.gidx = c->gidx.p;
    ta.g = c->gvec.p;
    ta.elr = c->elr.p;
    ta.counter = d_counter;
    return ta;
}

static SweepArgs sweep_rec_args(kmc_ctx* c) {
    SweepArgs sa;
    sa.g = c->geom();
    sa.lexp = c->lexp.p;
    sa.ncells = (int)c->ncells;
    sa.mode = 0;
    sa.rec = c->rec.p;
    sa.src_pos4 = nullptr;
    sa.src_cell = nullptr;
    sa.src_begin = sa.src_end = 0;
    return sa;
}

// work buffers sized for P moving charges and T proposals (a step allocates nothing)
static bool step_sized(const kmc_ctx* c, int64_t P, int64_t T) {
    const bool lr_sized =
        !c->have_lr || ((size_t)12 * P * lr_nbmax(c->L) <= c->aold.n &&
                        (size_t)12 * P * lr_split(P, nullptr) <= c->cold.n);
    return (size_t)(T + P) <= c->epos4.n && (size_t)T <= c->owner.n && (size_t)T <= c->rates.n &&
           (size_t)((T + 63) / 64) <= c->bsum.n && (size_t)(P + 1) <= c->es_copy.n && lr_sized;
}

static SelectArgs select_args(kmc_ctx* c, int64_t P, int64_t T, const int64_t* d_pid, double* d_dE, int64_t nsteps,
                              const double* d_uniform, double kT, int64_t* d_counter, int64_t* d_selected) {
    SelectArgs sa;
    sa.g = c->geom();
    sa.T = (int)T;
    sa.n = (int)c->n;
    sa.nsteps = nsteps;
    sa.kT = kT;
    sa.dE = d_dE;
    sa.rates = c->rates.p;
    sa.bsum = c->bsum.p;
    sa.uniform = d_uniform;
    sa.counter = d_counter;
    sa.selected = d_selected;
    sa.epos4 = c->epos4.p;
    sa.ecell = c->ecell.p;
    sa.owner = c->owner.p;
    sa.pid = d_pid;
    sa.ppos4 = c->ppos4.p;
    sa.pcell = c->pcell.p;
    sa.occ = c->occ.p;
    sa.occmax = c->occmax;
    sa.energy = c->energy.p;
    sa.rec = c->rec.p;
    sa.flags = c->flags.p;
    sa.err = c->err.p;
    (void)P;
    return sa;
}

// the side stream waits for everything issued so far on s
static int fork_side(kmc_ctx* c, hipStream_t s, int i) {
    HIPCHK(hipEventRecord(c->ev_fork[i], s));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork[i], 0));
    return KMC_OK;
}
// s waits for everything issued so far on the side stream
static int join_side(kmc_ctx* c, hipStream_t s, int i) {
    HIPCHK(hipEventRecord(c->ev_join[i], c->side));
    HIPCHK(hipStreamWaitEvent(s, c->ev_join[i], 0));
    return KMC_OK;
}

// ---- kernel profile (kmc_profile_*)
static bool prof_on(const kmc_ctx* c, hipStream_t s) {
    if (!c->prof) return false;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusNone;
}
static hipEvent_t prof_mark(kmc_ctx* c, hipStream_t s) {
    hipEvent_t e = nullptr;
    if (!c->prof_free.empty()) {
        e = c->prof_free.back();
        c->prof_free.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    if (hipEventRecord(e, s) != hipSuccess) {
        c->prof_free.push_back(e);
        return nullptr;
    }
    return e;
}
// waits for every pending pair and adds it to the sums
static int prof_collect(kmc_ctx* c) {
    for (auto& q : c->prof_pending) {
        HIPCHK(hipEventSynchronize(q.b));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, q.a, q.b));
        c->prof_ms[q.kind] += ms;
        c->prof_count[q.kind] += q.count;
        c->prof_free.push_back(q.a);
        c->prof_free.push_back(q.b);
    }
    c->prof_pending.clear();
    return KMC_OK;
}
static void prof_add(kmc_ctx* c, int kind, hipEvent_t a, hipEvent_t b, int64_t count) {
    if (a && b) c->prof_pending.push_back(kmc_ctx::ProfPair{a, b, kind, count});
    else {
        if (a) c->prof_free.push_back(a);
        if (b) c->prof_free.push_back(b);
    }
    if (c->prof_pending.size() >= 4096) prof_collect(c);   // (synchronises)
}

// w: the evaluation arrays (null: the context's own, sized here as needed for the long
// range); with w given nothing is allocated
static int run_propose_core(kmc_ctx* c, hipStream_t s, int64_t P, int64_t T, double* d_out, const int64_t* rate_loc,
                            bool mixed, double kT = 0.0, double* rates = nullptr, const int64_t* mask = nullptr,
                            int rate_rule = 0, int wait_sweep = 0, const ProposeWork* w = nullptr,
                            bool lr_first = false, bool far_side2 = false) {
    const int ne = (int)(T + P);
    if (!w) {   // sizes of the context's evaluation arrays (kmc_time_kernel reuses them)
        c->last_P = P;
        c->last_T = T;
    }
    const bool lr = (c->bc == KMC_BC_PBC || c->bc == KMC_BC_FF_ONLY) && c->have_lr;
    const bool prof = !w && prof_on(c, s);   // (the no-allocation path records no events)
    if (lr) {
        int rc = w ? KMC_OK : ensure_lr_work(c, P);
        if (rc) return rc;
        rc = fork_side(c, s, 0);
        if (rc) return rc;
    }
    // Plain steps record the near and far fields before the long range: a captured
    // graph's first branch after a fork stays on the launching stream's queue, and the
    // near/far chain is then the critical path (a dependency across queues costs
    // ~10 us; 2M charges: 0.767 -> 0.745 ms per step).  Pipelined runs (lr_first)
    // record the long range first, the order they were tuned with (100k charges, mode 2:
    // 0.114 against 0.121 ms the other way round).
    auto launch_lr = [&]() {
        hipEvent_t e0 = prof ? prof_mark(c, c->side) : nullptr;
        c->fns.lr(lr_args(c, P, T, w), c->side);
        if (prof) prof_add(c, 2, e0, prof_mark(c, c->side), T);
    };
    if (lr && lr_first) launch_lr();
    const ProposeWork wk = w ? *w : ctx_work(c);
    // far_side2 (pipelined mode 3): the far field on side2, behind the previous step's
    // sweep there (stream order) and after this step's set-up (event), so that it runs
    // beside the near field; s waits for it before the combine
    const bool far2 = far_side2 && c->bc != KMC_BC_FF_ONLY;
    if (far2) {
        HIPCHK(hipEventRecord(c->ev_setup, s));
        HIPCHK(hipStreamWaitEvent(c->side2, c->ev_setup, 0));
        EvalArgs af = eval_args(c, ne, wk.epos4, wk.ecell, wk.eexcl, wk.UI);
        hipEvent_t e1 = prof ? prof_mark(c, c->side2) : nullptr;
        c->fns.far(af, c->side2);
        if (prof) prof_add(c, 1, e1, prof_mark(c, c->side2), ne);
        HIPCHK(hipEventRecord(c->ev_far, c->side2));
    }
    if (c->bc != KMC_BC_FF_ONLY) {
        // wait_sweep 2: the near field too waits for the previous step's sweep (only the
        // set-up and the long range overlap it)
        if (wait_sweep == 2) HIPCHK(hipStreamWaitEvent(s, c->ev_sweep, 0));
        EvalArgs an = eval_args(c, ne, wk.epos4, wk.ecell, wk.eexcl, wk.UD);
        hipEvent_t e0 = prof ? prof_mark(c, s) : nullptr;
        launch_near_any(c, s, an, mixed, (int)P, (int)T, wk);
        if (prof) {
            prof_add(c, 0, e0, prof_mark(c, s), ne);
            if (c->prof_pairs.ensure(1) == hipSuccess && ne > 0)
                hipLaunchKernelGGL(k_pair_count, dim3((ne + 255) / 256), dim3(256), 0, s, an, c->prof_pairs.p);
        }
        if (far2) {
            HIPCHK(hipStreamWaitEvent(s, c->ev_far, 0));
        } else {
            // the far field reads the leaf expansions: after the previous step's sweep
            if (wait_sweep == 1) HIPCHK(hipStreamWaitEvent(s, c->ev_sweep, 0));
            EvalArgs af = eval_args(c, ne, wk.epos4, wk.ecell, wk.eexcl, wk.UI);
            hipEvent_t e1 = prof ? prof_mark(c, s) : nullptr;
            c->fns.far(af, s);
            if (prof) prof_add(c, 1, e1, prof_mark(c, s), ne);
        }
    }
    if (lr) {
        if (!lr_first) launch_lr();
        int rc = join_side(c, s, 0);
        if (rc) return rc;
    }
    CombineArgs ca = combine_args(c, P, T, d_out, rate_loc, kT, rates, w);
    ca.mask = mask;
    ca.rate_rule = rate_rule;
    if (T > 0) {
        if (mixed) hipLaunchKernelGGL(k_combine<true>, dim3((T + 63) / 64), dim3(64), 0, s, ca);
        else hipLaunchKernelGGL(k_combine<false>, dim3((T + 63) / 64), dim3(64), 0, s, ca);
    }
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

static int launch_setup(kmc_ctx* c, hipStream_t s, int64_t P, int64_t T, const int64_t* d_pid, const int64_t* d_es,
                        const double* d_new_pos, const double* d_new_q, const ProposeWork* w = nullptr,
                        const int64_t* d_mask = nullptr) {
    SetupArgs a = setup_args(c, P, T, d_pid, d_es, d_new_pos, d_new_q, w);
    a.mask = d_mask;
    const int64_t nthr = std::max(T, P);
    if (P > 0) hipLaunchKernelGGL(k_setup, dim3((nthr + 255) / 256), dim3(256), 0, s, a);
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

static int run_accept_tail(kmc_ctx* c, hipStream_t s, int64_t* d_counter, bool defer_sweep = false) {
    const TailArgs ta = tail_args(c, d_counter);
    if (c->bc == KMC_BC_FF_ONLY) {
        c->fns.md(ta, s);
        HIPCHK(hipGetLastError());
        return KMC_OK;
    }
    // the tail (cell tables, M, g, E_lr, counter) and the sweep (leaf expansions) only
    // share the move record, which both read: they run side by side
    int rc = fork_side(c, s, 1);
    if (rc) return rc;
    // plain steps record the sweep first (it stays on the launching queue), pipelined
    // steps the tail first (the order they were tuned with)
    if (defer_sweep) c->fns.md(ta, c->side);
    const SweepArgs sa = sweep_rec_args(c);
    const bool prof = prof_on(c, s);
    if (defer_sweep) {
        // the sweep on side2; s waits only for the tail, and the next step's far field
        // waits for ev_sweep (the caller joins ev_sweep into s after the last step)
        HIPCHK(hipEventRecord(c->ev_fork2, s));
        HIPCHK(hipStreamWaitEvent(c->side2, c->ev_fork2, 0));
        hipEvent_t e0 = prof ? prof_mark(c, c->side2) : nullptr;
        c->fns.sweep(sa, (int)c->ncells, c->side2, c->sweep_block, c->sweep_pairs, c->sweep_split);
        if (prof) prof_add(c, 3, e0, prof_mark(c, c->side2), 1);
        HIPCHK(hipEventRecord(c->ev_sweep, c->side2));
    } else {
        hipEvent_t e0 = prof ? prof_mark(c, s) : nullptr;
        c->fns.sweep(sa, (int)c->ncells, s, c->sweep_block, c->sweep_pairs, c->sweep_split);
        if (prof) prof_add(c, 3, e0, prof_mark(c, s), 1);
    }
    if (!defer_sweep) c->fns.md(ta, c->side);
    rc = join_side(c, s, 1);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

static int grow_store(kmc_ctx* c, int newmax, hipStream_t s) {
    if (newmax <= c->occmax) return KMC_OK;
    const size_t nslots = (size_t)c->ncells * newmax;
    double4* np = nullptr;
    int32_t* ni = nullptr;
    HIPCHK(hipMalloc((void**)&np, nslots * sizeof(double4)));
    HIPCHK(hipMalloc((void**)&ni, nslots * sizeof(int32_t)));
    HIPCHK(hipMemsetAsync(np, 0, nslots * sizeof(double4), s));
    HIPCHK(hipMemsetAsync(ni, 0xff, nslots * sizeof(int32_t), s));
    HIPCHK(hipMemcpy2DAsync(np, (size_t)newmax * sizeof(double4), c->slot_pos.p, (size_t)c->occmax * sizeof(double4),
                            (size_t)c->occmax * sizeof(double4), (size_t)c->ncells, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpy2DAsync(ni, (size_t)newmax * sizeof(int32_t), c->slot_id.p, (size_t)c->occmax * sizeof(int32_t),
                            (size_t)c->occmax * sizeof(int32_t), (size_t)c->ncells, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    c->slot_pos.release();
    c->slot_id.release();
    c->slot_pos.p = np;
    c->slot_pos.n = nslots;
    c->slot_id.p = ni;
    c->slot_id.n = nslots;
    c->occmax = newmax;
    c->generation++;
    return KMC_OK;
}

// grow a device array to hold `want` elements, keeping the first `used`
template <typename T>
static int grow_keep(kmc_ctx* c, DBuf<T>& b, size_t used, size_t want, hipStream_t s) {
    if (want <= b.n && b.p) return KMC_OK;
    const size_t cap = std::max(want, b.n + b.n / 2);
    T* np = nullptr;
    HIPCHK(hipMalloc((void**)&np, cap * sizeof(T)));
    if (used) HIPCHK(hipMemcpyAsync(np, b.p, used * sizeof(T), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    b.release();
    b.p = np;
    b.n = cap;
    c->generation++;
    return KMC_OK;
}

// host mirrors of positions, charges, leaves and occupancies from the device (after
// device-side steps moved charges)
static int refresh_mirrors(kmc_ctx* c, hipStream_t s) {
    if (!c->mirrors_stale) return KMC_OK;
    const int64_t n = c->n;
    std::vector<double4> p4(n);
    c->h_pos.resize(3 * n);
    c->h_q.resize(n);
    c->h_cell.resize(n);
    c->h_occ.resize(c->ncells);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(p4.data(), c->ppos4.p, n * sizeof(double4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(c->h_cell.data(), c->pcell.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(c->h_occ.data(), c->occ.p, c->ncells * sizeof(int), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
        c->h_pos[3 * i] = p4[i].x;
        c->h_pos[3 * i + 1] = p4[i].y;
        c->h_pos[3 * i + 2] = p4[i].z;
        c->h_q[i] = p4[i].w;
    }
    c->mirrors_stale = false;
    return KMC_OK;
}

// P2L of charges src[0..k) (charge in .w) into every well-separated leaf, with their
// 27 images for '27' / pbc, in source chunks
static void sweep_sources(kmc_ctx* c, hipStream_t s, const double4* src, const int32_t* src_cell, int k) {
    const int chunk = 2048;
    for (int b = 0; b < k; b += chunk) {
        SweepArgs sa;
        sa.g = c->geom();
        sa.lexp = c->lexp.p;
        sa.ncells = (int)c->ncells;
        sa.mode = 1;
        sa.rec = nullptr;
        sa.src_pos4 = src;
        sa.src_cell = src_cell;
        sa.src_begin = b;
        sa.src_end = std::min(k, b + chunk);
        c->fns.sweep(sa, (int)c->ncells, s, c->sweep_block, c->sweep_pairs, c->sweep_split);
    }
}

// M += multipole of charges src[0..k), then g = G M and E_lr (kmc_full_long_range.py:270-303)
static int lr_add_sources(kmc_ctx* c, hipStream_t s, const double4* src, int k) {
    const int chunk = 1024;
    const int nblk = (k + chunk - 1) / chunk;
    const int nrow = 2 * c->nh;
    DBuf<double> partial;
    HIPCHK(partial.ensure((size_t)nblk * nrow));
    c->fns.mp(k, chunk, src, partial.p, nblk, s);
    hipLaunchKernelGGL(k_reduce_rows_add, dim3((nrow + 255) / 256), dim3(256), 0, s, nblk, nrow, partial.p, c->M.p);
    hipLaunchKernelGGL(k_lr_refresh, dim3(1), dim3(256), 0, s, nrow, c->gdat.p, c->gptr.p, c->gidx.p, c->M.p,
                       c->gvec.p, c->elr.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return KMC_OK;
}

// eu-scaled energies of n evaluation entries: near field (without the excluded id),
// leaf expansion and long-range term.  The caller synchronises before the scratch
// buffers go out of scope.
static int entry_energies(kmc_ctx* c, hipStream_t s, int n, const double4* pos4, const int32_t* cell,
                          const int32_t* excl, double* out, DBuf<double>& ud, DBuf<double>& ui, DBuf<double>& lr,
                          double eu = -1.0) {
    if (eu < 0.0) eu = c->eu;
    const int nfar = far_parts(c->L);
    const bool near_far = c->bc != KMC_BC_FF_ONLY;
    const bool lrt = uses_lr(c);
    if (near_far) {
        HIPCHK(ud.ensure(n));
        HIPCHK(ui.ensure((size_t)n * nfar));
        launch_near(eval_args(c, n, pos4, cell, excl, ud.p), c->near_lanes, false, s);
        c->fns.far(eval_args(c, n, pos4, cell, excl, ui.p), s);
    }
    if (lrt) {
        HIPCHK(lr.ensure(n));
        hipLaunchKernelGGL(k_eval_lr, dim3((n + 255) / 256), dim3(256), 0, s, c->L, n, pos4, c->gvec.p, lr.p);
    }
    hipLaunchKernelGGL(k_charge_energy, dim3((n + 255) / 256), dim3(256), 0, s, n, pos4, near_far ? ud.p : nullptr,
                       near_far ? ui.p : nullptr, nfar, lrt ? lr.p : nullptr, eu, out);
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

static void launch_group_energy(kmc_ctx* c, hipStream_t s, int ngroups, int m, const double4* pos4, double* out) {
    hipLaunchKernelGGL(k_group_energy, dim3(ngroups), dim3(256), (size_t)2 * c->nh * sizeof(double), s, m, pos4, c->E,
                       c->bc, c->L, c->nh, uses_lr(c) ? c->gdat.p : nullptr, c->gptr.p, c->gidx.p, c->eu, out);
}

// propose_extract of nrows rows of m charges given by device ids (synchronises)
static int propose_extract_core(kmc_ctx* c, hipStream_t s, int nrows, int m, const int64_t* d_ids, bool self_energy,
                                double* d_out) {
    const int n = nrows * m;
    DBuf<double4> pos4;
    DBuf<int32_t> cell, excl;
    DBuf<double> ce, ge, ud, ui, lr;
    HIPCHK(pos4.ensure(n));
    HIPCHK(cell.ensure(n));
    HIPCHK(excl.ensure(n));
    HIPCHK(ce.ensure(n));
    HIPCHK(ge.ensure(nrows));
    HIPCHK(hipMemsetAsync(c->err.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_gather_ids, dim3((n + 255) / 256), dim3(256), 0, s, c->geom(), n, d_ids, c->n, c->ppos4.p,
                       c->pcell.p, 1.0, pos4.p, cell.p, excl.p, c->err.p);
    int rc = entry_energies(c, s, n, pos4.p, cell.p, excl.p, ce.p, ud, ui, lr);
    if (rc) return rc;
    if (self_energy) launch_group_energy(c, s, nrows, m, pos4.p, ge.p);
    hipLaunchKernelGGL(k_extract_rows, dim3((nrows + 255) / 256), dim3(256), 0, s, nrows, m,
                       self_energy ? ge.p : nullptr, ce.p, d_out);
    HIPCHK(hipGetLastError());
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, c->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail_err_counter(c, s, KMC_E_ARG, "charge id out of range");
    return KMC_OK;
}

// propose_inject of n charges (device positions [n][3], charges [n]) into d_out[0]
// (synchronises)
static int propose_inject_core(kmc_ctx* c, hipStream_t s, int n, const double* d_pos, const double* d_q, double* d_out) {
    DBuf<double4> pos4;
    DBuf<int32_t> cell, excl;
    DBuf<double> ce, ge, ud, ui, lr;
    HIPCHK(pos4.ensure(n));
    HIPCHK(cell.ensure(n));
    HIPCHK(excl.ensure(n));
    HIPCHK(ce.ensure(n));
    HIPCHK(ge.ensure(1));
    hipLaunchKernelGGL(k_pack_charges, dim3((n + 255) / 256), dim3(256), 0, s, c->geom(), n, d_pos, d_q, pos4.p, cell.p,
                       excl.p);
    int rc = entry_energies(c, s, n, pos4.p, cell.p, excl.p, ce.p, ud, ui, lr);
    if (rc) return rc;
    launch_group_energy(c, s, 1, n, pos4.p, ge.p);
    hipLaunchKernelGGL(k_inject_sum, dim3(1), dim3(256), 0, s, n, ce.p, ge.p, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return KMC_OK;
}

// ============================================================================
// C ABI
// ============================================================================
// ---------------------------------------------------------------------------
// Fixed scratch of kmc_propose_noalloc: the evaluation arrays of one propose carved
// from one file-scope device array of KMC_NOALLOC_SCRATCH_BYTES.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code never uses
// the array's name as an address: the device address is taken once per device, with
// that device current, by hipGetSymbolAddress, and only that pointer reaches kernels.
// ---------------------------------------------------------------------------
constexpr int KMC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[KMC_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[KMC_MAX_DEVICES] = {};

// the current device's copy of the scratch array
static int scratch_ptr(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= KMC_MAX_DEVICES) return set_err(KMC_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err(KMC_E_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return KMC_OK;
}

// Bytes of scratch a propose of P moving charges and T proposals takes, and (base !=
// null) the carve of the evaluation arrays from base, each 256-byte aligned.
static size_t carve_work(const kmc_ctx* c, int64_t P, int64_t T, unsigned char* base, ProposeWork* w) {
    const size_t ne = (size_t)(T + P), nfar = (size_t)far_parts(c->L);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return (void*)p;
    };
    ProposeWork k{};
    k.err = (int*)take(sizeof(int));
    k.epos4 = (double4*)take(ne * sizeof(double4));
    k.ecell = (int32_t*)take(ne * sizeof(int32_t));
    k.eexcl = (int32_t*)take(ne * sizeof(int32_t));
    k.owner = (int32_t*)take((size_t)T * sizeof(int32_t));
    k.UD = (double*)take(ne * sizeof(double));
    k.UI = (double*)take(ne * nfar * sizeof(double));
    k.es_copy = (int64_t*)take((size_t)(P + 1) * sizeof(int64_t));
    const int nt = near_tiles(c);
    if (nt > 0) {
        k.tcount = (int*)take((size_t)nt * TSTRIDE * sizeof(int));
        k.toff = (int*)take((size_t)(nt + 1) * sizeof(int));
        k.rank = (int*)take(ne * sizeof(int));
        k.perm = (int*)take(ne * sizeof(int));
    }
    if (uses_lr(c)) {
        k.lrd = (double*)take((size_t)12 * T * sizeof(double));
        k.aold = (double*)take((size_t)12 * P * lr_nbmax(c->L) * sizeof(double));
        k.cold = (double*)take((size_t)12 * P * lr_split(P) * sizeof(double));
    }
    if (w) *w = k;
    return off;
}

extern "C" {

const char* kmc_last_error(void) { return g_err.c_str(); }
int32_t kmc_min_order(void) { return KMIN_L; }
int32_t kmc_max_order(void) { return KMAX_L; }

int kmc_create(kmc_ctx** out, int64_t n, double extent, int32_t r, int32_t L, int32_t bc, double max_move,
               double energy_unit, int32_t occ_headroom) {
    if (!out) return set_err(KMC_E_ARG, "null output pointer");
    *out = nullptr;
    if (n < 1 || n >= (int64_t)1 << 31) return set_err(KMC_E_ARG, "n out of range");
    if (!(extent > 0)) return set_err(KMC_E_ARG, "extent must be positive");
    if (r < 1 || r > 11) return set_err(KMC_E_ARG, "r out of range");
    if (bc < 0 || bc > 3) return set_err(KMC_E_ARG, "unknown boundary condition");
    kmc_ctx* c = new kmc_ctx();
    if (!order_fns(L, &c->fns)) {
        delete c;
        return set_err(KMC_E_UNSUPPORTED, "expansion order L=" + std::to_string(L) + " not built (2..20)");
    }
    c->n = n;
    c->E = extent;
    c->r = r;
    c->L = L;
    c->bc = bc;
    c->ncps = 1 << (r - 1);
    c->ncells = (int64_t)c->ncps * c->ncps * c->ncps;
    c->w = extent / c->ncps;
    c->nh = nhalf(L);
    c->max_move = max_move > 0 ? max_move : extent;
    c->eu = energy_unit;
    c->occ_headroom = occ_headroom >= 0 ? occ_headroom : 8;
    // kmc_fmm_self_interaction.py:48-54
    double s27 = 0.0;
    for (int k = 0; k < 27; k++) {
        const int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
        if (ox == 0 && oy == 0 && oz == 0) continue;
        const double dx = extent * ox, dy = extent * oy, dz = extent * oz;
        s27 -= 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    }
    c->self27 = s27;
    int rc = KMC_OK;
    if (c->energy.ensure(1) != hipSuccess || c->flags.ensure(1) != hipSuccess || c->err.ensure(1) != hipSuccess ||
        c->rec.ensure(1) != hipSuccess || c->elr.ensure(1) != hipSuccess || c->dE1.ensure(1) != hipSuccess ||
        c->one_pid.ensure(1) != hipSuccess || c->one_es.ensure(2) != hipSuccess || c->one_pos.ensure(3) != hipSuccess) {
        rc = set_err(KMC_E_HIP, "device allocation failed");
    }
    if (rc == KMC_OK) {
        if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_sweep, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_setup, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_far, hipEventDisableTiming) != hipSuccess)
            rc = set_err(KMC_E_HIP, "stream");
        for (int i = 0; i < 2 && rc == KMC_OK; i++)
            if (hipEventCreateWithFlags(&c->ev_fork[i], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming) != hipSuccess)
                rc = set_err(KMC_E_HIP, "event");
    }
    if (rc == KMC_OK) {
        hipMemset(c->energy.p, 0, sizeof(double));
        hipMemset(c->flags.p, 0, sizeof(int));
        hipMemset(c->err.p, 0, sizeof(int));
        hipMemset(c->elr.p, 0, sizeof(double));
        MoveRec z{};
        hipMemcpy(c->rec.p, &z, sizeof(MoveRec), hipMemcpyHostToDevice);
    }
    if (rc != KMC_OK) {
        kmc_destroy(c);
        return rc;
    }
    *out = c;
    return KMC_OK;
}

int kmc_destroy(kmc_ctx* c) {
    if (!c) return KMC_OK;
    c->ppos4.release(); c->pcell.release(); c->occ.release(); c->slot_pos.release(); c->slot_id.release();
    c->lexp.release(); c->M.release(); c->gvec.release(); c->gdat.release(); c->elr.release(); c->gptr.release();
    c->gblk.release(); c->aold.release(); c->cold.release();
    c->gidx.release(); c->energy.release(); c->flags.release(); c->err.release(); c->rec.release();
    c->epos4.release(); c->ecell.release(); c->eexcl.release(); c->owner.release(); c->UD.release();
    c->UI.release(); c->lrd.release(); c->dE1.release(); c->one_pid.release(); c->one_es.release();
    c->one_pos.release(); c->d_cnt.release(); c->d_has.release(); c->d_slot.release(); c->d_off.release();
    c->d_pid.release(); c->d_es.release(); c->d_rate.release(); c->d_tot.release(); c->d_npos.release();
    c->d_nq.release(); c->rates.release(); c->bsum.release();
    for (int i = 0; i < 2; i++) {
        if (c->ev_fork[i]) hipEventDestroy(c->ev_fork[i]);
        if (c->ev_join[i]) hipEventDestroy(c->ev_join[i]);
    }
    if (c->side) hipStreamDestroy(c->side);
    if (c->ev_fork2) hipEventDestroy(c->ev_fork2);
    if (c->ev_sweep) hipEventDestroy(c->ev_sweep);
    if (c->ev_setup) hipEventDestroy(c->ev_setup);
    if (c->ev_far) hipEventDestroy(c->ev_far);
    if (c->side2) hipStreamDestroy(c->side2);
    for (auto& q : c->prof_pending) {
        hipEventDestroy(q.a);
        hipEventDestroy(q.b);
    }
    for (hipEvent_t e : c->prof_free) hipEventDestroy(e);
    delete c;
    return KMC_OK;
}

int kmc_set_tiles(kmc_ctx* c, int32_t near_lanes, int32_t sweep_block, int32_t sweep_pairs, int32_t sweep_split) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (near_lanes != 0 && near_lanes != 1 && near_lanes != 4 && near_lanes != 8 && near_lanes != 16 &&
        near_lanes != 27 && near_lanes != 32 && near_lanes != 64 && near_lanes != -4 && near_lanes != -8 &&
        near_lanes != -16)
        return set_err(KMC_E_ARG, "near_lanes must be one of 0,1,4,8,16,27,32,64,-4,-8,-16 (27: one lane per "
                                  "neighbour leaf, 0: the proposals of a moving charge together, -G: points sorted "
                                  "by 4x4x4 leaf tile, the tile's charges staged in LDS, G lanes per point)");
    if (sweep_pairs < 1 || sweep_pairs > 2) return set_err(KMC_E_ARG, "sweep_pairs must be 1 or 2");
    if (sweep_block != 0 && sweep_block != 1 && (sweep_block < 64 || sweep_block > 256 || sweep_block % 64))
        return set_err(KMC_E_ARG, "sweep_block must be 0 (fused), 1 (4x4x4 leaf blocks with order truncation), "
                                  "64, 128, 192 or 256");
    if (sweep_split != 1 && sweep_split != 3) return set_err(KMC_E_ARG, "sweep_split must be 1 or 3");
    c->near_lanes = near_lanes;
    c->sweep_block = sweep_block;
    c->sweep_pairs = sweep_pairs;
    c->sweep_split = sweep_split;
    return KMC_OK;
}

int kmc_profile_enable(kmc_ctx* c, int32_t on) {
    if (check_ctx(c)) return KMC_E_ARG;
    int rc = prof_collect(c);
    if (rc) return rc;
    for (int k = 0; k < 4; k++) {
        c->prof_ms[k] = 0.0;
        c->prof_count[k] = 0;
    }
    HIPCHK(c->prof_pairs.ensure(1));
    HIPCHK(hipMemset(c->prof_pairs.p, 0, sizeof(unsigned long long)));
    c->prof = on != 0;
    return KMC_OK;
}

int kmc_profile_read(kmc_ctx* c, double* h_out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!h_out) return set_err(KMC_E_ARG, "null output");
    int rc = prof_collect(c);
    if (rc) return rc;
    unsigned long long pairs = 0;
    if (c->prof_pairs.p) {
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(&pairs, c->prof_pairs.p, sizeof(pairs), hipMemcpyDeviceToHost));
    }
    h_out[0] = c->prof_ms[0];
    h_out[1] = (double)c->prof_count[0];
    h_out[2] = (double)pairs;
    h_out[3] = c->prof_ms[1];
    h_out[4] = (double)c->prof_count[1];
    h_out[5] = c->prof_ms[2];
    h_out[6] = (double)c->prof_count[2];
    h_out[7] = c->prof_ms[3];
    h_out[8] = (double)c->prof_count[3];
    return KMC_OK;
}

int kmc_set_near_order(kmc_ctx* c, int32_t order) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (order != 0 && order != 1) return set_err(KMC_E_ARG, "near order must be 0 (proposal order) or 1 (leaf tiles)");
    c->near_order = order;
    return KMC_OK;
}

int kmc_set_particles(kmc_ctx* c, const double* h_pos, const double* h_q) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!h_pos || !h_q) return set_err(KMC_E_ARG, "null input");
    const int64_t n = c->n;
    c->h_pos.assign(h_pos, h_pos + 3 * n);
    c->h_q.assign(h_q, h_q + n);
    c->h_cell.resize(n);
    c->h_occ.assign(c->ncells, 0);
    for (int64_t i = 0; i < n; i++) {
        const int cx = bin_axis(h_pos[3 * i], c->E, c->w, c->ncps);
        const int cy = bin_axis(h_pos[3 * i + 1], c->E, c->w, c->ncps);
        const int cz = bin_axis(h_pos[3 * i + 2], c->E, c->w, c->ncps);
        c->h_cell[i] = cx + c->ncps * (cy + c->ncps * cz);
        c->h_occ[c->h_cell[i]]++;
    }
    int mx = 1;
    for (int v : c->h_occ) mx = std::max(mx, v);
    c->occmax = mx + c->occ_headroom;
    // particles of a leaf in ascending index (kmc_local.py:606-615)
    std::vector<int64_t> start(c->ncells + 1, 0);
    for (int64_t k = 0; k < c->ncells; k++) start[k + 1] = start[k] + c->h_occ[k];
    std::vector<double4> slots((size_t)c->ncells * c->occmax, make_double4(0, 0, 0, 0));
    std::vector<int32_t> ids((size_t)c->ncells * c->occmax, -1);
    std::vector<int> fill(c->ncells, 0);
    std::vector<double4> p4(n);
    for (int64_t i = 0; i < n; i++) {
        const int k = c->h_cell[i];
        const size_t sidx = (size_t)k * c->occmax + fill[k]++;
        p4[i] = make_double4(h_pos[3 * i], h_pos[3 * i + 1], h_pos[3 * i + 2], h_q[i]);
        slots[sidx] = p4[i];
        ids[sidx] = (int32_t)i;
    }
    HIPCHK(c->ppos4.ensure(n));
    HIPCHK(c->pcell.ensure(n));
    HIPCHK(c->occ.ensure(c->ncells));
    HIPCHK(c->slot_pos.ensure(slots.size()));
    HIPCHK(c->slot_id.ensure(ids.size()));
    HIPCHK(c->lexp.ensure((size_t)c->ncells * c->nh));
    HIPCHK(hipMemcpy(c->ppos4.p, p4.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->pcell.p, c->h_cell.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->occ.p, c->h_occ.data(), c->ncells * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->slot_pos.p, slots.data(), slots.size() * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->slot_id.p, ids.data(), ids.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->lexp.p, 0, (size_t)c->ncells * c->nh * sizeof(double2)));
    HIPCHK(c->M.ensure(2 * c->nh));
    HIPCHK(c->gvec.ensure(2 * c->nh));
    HIPCHK(hipMemset(c->M.p, 0, 2 * c->nh * sizeof(double)));
    HIPCHK(hipMemset(c->gvec.p, 0, 2 * c->nh * sizeof(double)));
    c->have_particles = true;
    c->initialised = false;
    c->mirrors_stale = false;
    c->generation++;
    return KMC_OK;
}

int kmc_set_lr_operator(kmc_ctx* c, int64_t nnz, const double* h_data, const int64_t* h_indptr,
                        const int64_t* h_indices) {
    if (check_ctx(c)) return KMC_E_ARG;
    const int nrow = 2 * c->nh;
    if (nnz < 0 || !h_indptr || (nnz > 0 && (!h_data || !h_indices))) return set_err(KMC_E_ARG, "bad CSR");
    if (h_indptr[0] != 0 || h_indptr[nrow] != nnz) return set_err(KMC_E_ARG, "CSR indptr does not match 2*L(L+1)/2 rows");
    std::vector<int32_t> idx(std::max<int64_t>(nnz, 1), 0);
    for (int64_t k = 0; k < nnz; k++) {
        if (h_indices[k] < 0 || h_indices[k] >= nrow) return set_err(KMC_E_ARG, "CSR column out of range");
        idx[k] = (int32_t)h_indices[k];
    }
    HIPCHK(c->gdat.ensure(std::max<int64_t>(nnz, 1)));
    HIPCHK(c->gidx.ensure(std::max<int64_t>(nnz, 1)));
    HIPCHK(c->gptr.ensure(nrow + 1));
    if (nnz > 0) {
        HIPCHK(hipMemcpy(c->gdat.p, h_data, nnz * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->gidx.p, idx.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->gptr.p, h_indptr, (nrow + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    // dense diagonal blocks in the kernels' order; everything outside them must vanish
    {
        std::vector<double> dense((size_t)nrow * nrow, 0.0);
        for (int r = 0; r < nrow; r++)
            for (int64_t k = h_indptr[r]; k < h_indptr[r + 1]; k++) dense[(size_t)r * nrow + h_indices[k]] += h_data[k];
        const LRBlocks T = make_lr_blocks(c->L);
        std::vector<int> blk(nrow, -1);
        size_t total = 0;
        for (int b = 0; b < 6; b++) {
            for (int i = 0; i < T.n[b]; i++)
                for (int comp = 0; comp < 2; comp++) blk[2 * hidx(T.l[b][i], T.m[b][i]) + comp] = 2 * b + comp;
            total += 2 * (size_t)T.n[b] * T.n[b];
        }
        for (int r = 0; r < nrow; r++)
            for (int q = 0; q < nrow; q++)
                if (dense[(size_t)r * nrow + q] != 0.0 && blk[r] != blk[q])
                    return set_err(KMC_E_ARG, "long-range operator couples entries outside the cubic-symmetry blocks");
        std::vector<double> gb(std::max<size_t>(total, 1), 0.0);
        for (int b = 0; b < 6; b++) {
            const int nb = T.n[b];
            for (int comp = 0; comp < 2; comp++)
                for (int i = 0; i < nb; i++)
                    for (int j = 0; j < nb; j++) {
                        const int ri = 2 * hidx(T.l[b][i], T.m[b][i]) + comp;
                        const int rj = 2 * hidx(T.l[b][j], T.m[b][j]) + comp;
                        gb[T.off[b] + (size_t)comp * nb * nb + (size_t)i * nb + j] = dense[(size_t)ri * nrow + rj];
                    }
        }
        HIPCHK(c->gblk.ensure(gb.size()));
        HIPCHK(hipMemcpy(c->gblk.p, gb.data(), gb.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    c->gnnz = nnz;
    c->have_lr = true;
    return KMC_OK;
}

// Hierarchical initial solve of the leaf expansions (see the kernels above).
static bool use_hierarchical(const kmc_ctx* c) {
    if (c->bc == KMC_BC_FF_ONLY || c->ncps < 8) return false;
    if (c->init_method == 1) return false;
    if (c->init_method == 2) return true;
    return (double)c->n * (double)c->ncells > 2.0e10;
}

static int run_fmm_init(kmc_ctx* c, hipStream_t s) {
    // translations run at an internal order Li > L (M2L truncation error falls
    // geometrically with the order); the leaves receive orders l < L
    const int L = std::min(c->L + 4, KMAX_L), nh = nhalf(L);
    const int Lleaf = c->L;
    int R = 0;
    while ((1 << R) < c->ncps) R++;
    const bool periodic = c->bc != KMC_BC_FREE_SPACE;
    const int lo = periodic ? 1 : 2, hi = R - 2;
    // tables: N_pq for p <= 2L-2; conj(R~) (M2M) and R~ (L2L) at the 8 child offsets
    const int P = 2 * L - 1;
    std::vector<double> nrm(nhalf(P));
    for (int p = 0; p < P; p++)
        for (int q = 0; q <= p; q++) nrm[hidx(p, q)] = exp(0.5 * (lgamma(p - q + 1.0) + lgamma(p + q + 1.0)));
    const int TW = 2 * L - 1;
    std::vector<double2> rm((size_t)8 * L * TW, make_double2(0.0, 0.0)), rl(rm.size(), make_double2(0.0, 0.0));
    for (int oct = 0; oct < 8; oct++) {
        const double bx = (oct & 1) - 0.5, by = ((oct >> 1) & 1) - 0.5, bz = (oct >> 2) - 0.5;
        for (int nn = 0; nn < L; nn++)
            for (int mm = 0; mm <= nn; mm++) {
                const Cplx v = solid_lm<false>(nn, mm, bx, by, bz);
                const double iv = 1.0 / nrm[hidx(nn, mm)];
                const double2 r = make_double2(v.re * iv, v.im * iv);               // R~_n^m
                const double sg = (mm & 1) ? -1.0 : 1.0;
                const double2 rneg = make_double2(sg * r.x, -sg * r.y);           // R~_n^-m
                const size_t base = ((size_t)oct * L + nn) * TW + L - 1;
                rl[base + mm] = r;
                rl[base - mm] = rneg;
                rm[base + mm] = make_double2(r.x, -r.y);
                rm[base - mm] = make_double2(rneg.x, -rneg.y);
            }
    }
    DBuf<double> d_nrm;
    DBuf<double2> d_rm, d_rl;
    HIPCHK(d_nrm.ensure(nrm.size()));
    HIPCHK(d_rm.ensure(rm.size()));
    HIPCHK(d_rl.ensure(rl.size()));
    HIPCHK(hipMemcpy(d_nrm.p, nrm.data(), nrm.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_rm.p, rm.data(), rm.size() * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_rl.p, rl.data(), rl.size() * sizeof(double2), hipMemcpyHostToDevice));
    const Geom g = c->geom();
    auto cells = [](int lev) { return (long long)1 << (3 * lev); };
    auto grid = [nh](long long ncell) { return dim3((unsigned)((ncell * nh + 255) / 256)); };
    std::vector<DBuf<double2>> Q(R + 1), Lv(R + 1);
    if (hi >= lo) {
        HIPCHK(Q[hi].ensure((size_t)cells(hi) * nh));
        FmmLevel lv{1 << hi, c->E / (1 << hi)};
        hipLaunchKernelGGL(k_fmm_p2m, grid(cells(hi)), dim3(256), 0, s, L, lv, g, c->occ.p, c->slot_pos.p, c->occmax,
                           d_nrm.p, Q[hi].p);
        for (int lev = hi - 1; lev >= lo; lev--) {
            HIPCHK(Q[lev].ensure((size_t)cells(lev) * nh));
            hipLaunchKernelGGL(k_fmm_m2m, grid(cells(lev)), dim3(256), 0, s, L, 1 << lev, Q[lev + 1].p, d_rm.p,
                               Q[lev].p);
        }
        const int m2l_threads = std::max(64, ((std::max(nh, P) + 63) / 64) * 64);
        const size_t m2l_lds = ((size_t)P * (4 * L - 3) + nh) * sizeof(double2);
        for (int lev = lo; lev <= hi; lev++) {
            HIPCHK(Lv[lev].ensure((size_t)cells(lev) * nh));
            if (lev == lo) {
                HIPCHK(hipMemsetAsync(Lv[lev].p, 0, (size_t)cells(lev) * nh * sizeof(double2), s));
            } else {
                hipLaunchKernelGGL(k_fmm_l2l, grid(cells(lev)), dim3(256), 0, s, L, L, 1 << lev, Lv[lev - 1].p,
                                   d_rl.p, Lv[lev].p, 0, 0.0, d_nrm.p);
            }
            hipLaunchKernelGGL(k_fmm_m2l, dim3((unsigned)cells(lev)), dim3(m2l_threads), m2l_lds, s, L, 1 << lev,
                               periodic ? 1 : 0, Q[lev].p, d_nrm.p, Lv[lev].p);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(Lv[R - 1].ensure((size_t)cells(R - 1) * nh));
        hipLaunchKernelGGL(k_fmm_l2l, grid(cells(R - 1)), dim3(256), 0, s, L, L, 1 << (R - 1), Lv[hi].p, d_rl.p,
                           Lv[R - 1].p, 0, 0.0, d_nrm.p);
        const long long nleaf_out = cells(R) * c->nh;
        hipLaunchKernelGGL(k_fmm_l2l, dim3((unsigned)((nleaf_out + 255) / 256)), dim3(256), 0, s, L, Lleaf, 1 << R,
                           Lv[R - 1].p, d_rl.p, c->lexp.p, 1, c->w, d_nrm.p);
        HIPCHK(hipGetLastError());
    }
    SweepArgs sa;
    sa.g = g;
    sa.lexp = c->lexp.p;
    sa.ncells = (int)c->ncells;
    sa.mode = 1;
    sa.rec = nullptr;
    sa.src_pos4 = nullptr;
    sa.src_cell = nullptr;
    sa.src_begin = sa.src_end = 0;
    c->fns.p2ln(sa, c->occ.p, c->slot_pos.p, c->occmax, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return KMC_OK;
}

int kmc_set_init_method(kmc_ctx* c, int32_t method) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (method < 0 || method > 2) return set_err(KMC_E_ARG, "init method must be 0 (auto), 1 (direct) or 2 (hierarchical)");
    c->init_method = method;
    return KMC_OK;
}

int kmc_initialise(kmc_ctx* c, void* stream, double* h_energy) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles before initialise");
    if ((c->bc == KMC_BC_PBC || c->bc == KMC_BC_FF_ONLY) && !c->have_lr)
        return set_err(KMC_E_STATE, "set the long-range operator before initialise (pbc / ff-only)");
    hipStream_t s = (hipStream_t)stream;
    const int n = (int)c->n;
    HIPCHK(hipMemsetAsync(c->lexp.p, 0, (size_t)c->ncells * c->nh * sizeof(double2), s));
    if (use_hierarchical(c)) {
        int rc = run_fmm_init(c, s);
        if (rc) return rc;
    } else if (c->bc != KMC_BC_FF_ONLY) {
        // P2L of every charge into every well-separated leaf
        sweep_sources(c, s, c->ppos4.p, c->pcell.p, n);
        HIPCHK(hipGetLastError());
    }
    // level-0 multipole M
    {
        const int chunk = 1024;
        const int nblk = (n + chunk - 1) / chunk;
        DBuf<double> partial;
        HIPCHK(partial.ensure((size_t)nblk * 2 * c->nh));
        c->fns.mp(n, chunk, c->ppos4.p, partial.p, nblk, s);
        hipLaunchKernelGGL(k_reduce_rows, dim3((2 * c->nh + 255) / 256), dim3(256), 0, s, nblk, 2 * c->nh, partial.p,
                           c->M.p);
        HIPCHK(hipStreamSynchronize(s));
        partial.release();
    }
    if (c->have_lr)
        hipLaunchKernelGGL(k_lr_refresh, dim3(1), dim3(256), 0, s, 2 * c->nh, c->gdat.p, c->gptr.p, c->gidx.p, c->M.p,
                           c->gvec.p, c->elr.p);
    // energy: 1/2 sum_i q_i phi(x_i) + long range
    DBuf<int32_t> excl;
    DBuf<double> ud, ui, esr;
    HIPCHK(excl.ensure(n));
    HIPCHK(ud.ensure(n));
    HIPCHK(ui.ensure((size_t)n * far_parts(c->L)));
    HIPCHK(esr.ensure(1));
    {
        std::vector<int32_t> ids(n);
        for (int i = 0; i < n; i++) ids[i] = i;
        HIPCHK(hipMemcpyAsync(excl.p, ids.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(esr.p, 0, sizeof(double), s));
        if (c->bc != KMC_BC_FF_ONLY) {
            EvalArgs an = eval_args(c, n, c->ppos4.p, c->pcell.p, excl.p, ud.p);
            launch_near(an, c->near_lanes, false, s);
            EvalArgs af = eval_args(c, n, c->ppos4.p, c->pcell.p, excl.p, ui.p);
            c->fns.far(af, s);
            hipLaunchKernelGGL(k_half_sum, dim3(1), dim3(256), 0, s, n, ud.p, ui.p, far_parts(c->L), esr.p);
        }
        hipLaunchKernelGGL(k_finish_energy, dim3(1), dim3(1), 0, s, esr.p, c->elr.p, c->eu, c->bc, c->energy.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }
    excl.release();
    ud.release();
    ui.release();
    esr.release();
    c->initialised = true;
    if (h_energy) HIPCHK(hipMemcpy(h_energy, c->energy.p, sizeof(double), hipMemcpyDeviceToHost));
    return KMC_OK;
}

int kmc_get_expansions(kmc_ctx* c, double* h_out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first (kmc_set_particles)");
    const int L = c->L, L2 = L * L;
    std::vector<double2> h((size_t)c->ncells * c->nh);
    HIPCHK(hipMemcpy(h.data(), c->lexp.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < c->ncells; k++) {
        double* o = h_out + (size_t)k * 2 * L2;
        const double2* v = &h[(size_t)k * c->nh];
        for (int l = 0; l < L; l++)
            for (int m = -l; m <= l; m++) {
                const double2 x = v[hidx(l, m < 0 ? -m : m)];
                o[l * l + l + m] = x.x;
                o[L2 + l * l + l + m] = m < 0 ? -x.y : x.y;
            }
    }
    return KMC_OK;
}

int kmc_set_expansions(kmc_ctx* c, const double* h_in) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first");
    const int L = c->L, L2 = L * L;
    std::vector<double2> h((size_t)c->ncells * c->nh);
    for (int64_t k = 0; k < c->ncells; k++) {
        const double* o = h_in + (size_t)k * 2 * L2;
        for (int l = 0; l < L; l++)
            for (int m = 0; m <= l; m++) h[(size_t)k * c->nh + hidx(l, m)] = make_double2(o[l * l + l + m], o[L2 + l * l + l + m]);
    }
    HIPCHK(hipMemcpy(c->lexp.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
    return KMC_OK;
}

int kmc_get_multipole(kmc_ctx* c, double* h_out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    const int L = c->L, L2 = L * L;
    std::vector<double> h(2 * c->nh);
    HIPCHK(hipMemcpy(h.data(), c->M.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int l = 0; l < L; l++)
        for (int m = -l; m <= l; m++) {
            const int hh = hidx(l, m < 0 ? -m : m);
            h_out[l * l + l + m] = h[2 * hh];
            h_out[L2 + l * l + l + m] = m < 0 ? -h[2 * hh + 1] : h[2 * hh + 1];
        }
    return KMC_OK;
}

int kmc_store_occmax(kmc_ctx* c, int32_t* out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first (kmc_set_particles)");
    *out = c->occmax;
    return KMC_OK;
}

int kmc_get_store(kmc_ctx* c, int32_t* h_occ, int32_t* h_ids, double* h_pos4, int32_t* h_occmax) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first (kmc_set_particles)");
    if (h_occmax) *h_occmax = c->occmax;
    if (h_occ) HIPCHK(hipMemcpy(h_occ, c->occ.p, c->ncells * sizeof(int), hipMemcpyDeviceToHost));
    if (h_ids) HIPCHK(hipMemcpy(h_ids, c->slot_id.p, (size_t)c->ncells * c->occmax * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (h_pos4) HIPCHK(hipMemcpy(h_pos4, c->slot_pos.p, (size_t)c->ncells * c->occmax * sizeof(double4), hipMemcpyDeviceToHost));
    return KMC_OK;
}

int kmc_get_particles(kmc_ctx* c, double* h_pos, int32_t* h_cells) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first (kmc_set_particles)");
    std::vector<double4> p4(c->n);
    HIPCHK(hipMemcpy(p4.data(), c->ppos4.p, c->n * sizeof(double4), hipMemcpyDeviceToHost));
    if (h_pos)
        for (int64_t i = 0; i < c->n; i++) {
            h_pos[3 * i] = p4[i].x;
            h_pos[3 * i + 1] = p4[i].y;
            h_pos[3 * i + 2] = p4[i].z;
        }
    if (h_cells) HIPCHK(hipMemcpy(h_cells, c->pcell.p, c->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return KMC_OK;
}

int kmc_get_energy(kmc_ctx* c, double* h_energy) {
    if (check_ctx(c)) return KMC_E_ARG;
    HIPCHK(hipMemcpy(h_energy, c->energy.p, sizeof(double), hipMemcpyDeviceToHost));
    return KMC_OK;
}

int kmc_set_energy(kmc_ctx* c, double e) {
    if (check_ctx(c)) return KMC_E_ARG;
    HIPCHK(hipMemcpy(c->energy.p, &e, sizeof(double), hipMemcpyHostToDevice));
    return KMC_OK;
}

int kmc_ncells(kmc_ctx* c, int64_t* out) {
    if (check_ctx(c)) return KMC_E_ARG;
    *out = c->ncells;
    return KMC_OK;
}

int kmc_propose(kmc_ctx* c, void* stream, int64_t P, int64_t T, const int64_t* d_pid, const int64_t* d_es,
                const double* d_new_pos, const double* d_new_q, double* d_dE, int32_t precision) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (P < 0 || T < 0 || T >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad proposal counts");
    if (P == 0 || T == 0) return KMC_OK;
    if (!d_pid || !d_es || !d_new_pos || !d_dE) return set_err(KMC_E_ARG, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    int rc = ensure_eval(c, T, P);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(c->err.p, 0, sizeof(int), s));
    rc = launch_setup(c, s, P, T, d_pid, d_es, d_new_pos, d_new_q);
    if (rc) return rc;
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, c->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail_err_counter(c, s, KMC_E_BAD_MOVE, "Bad proposed move was detected.");
    return run_propose_core(c, s, P, T, d_dE, nullptr, precision == KMC_PREC_MIXED);
}

int kmc_propose_noalloc(kmc_ctx* c, void* stream, int64_t P, int64_t T, const int64_t* d_pid, const int64_t* d_es,
                        const double* d_new_pos, const double* d_new_q, double* d_dE, int32_t precision) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (P < 0 || T < 0 || T >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad proposal counts");
    if (P == 0 || T == 0) return KMC_OK;
    if (!d_pid || !d_es || !d_new_pos || !d_dE) return set_err(KMC_E_ARG, "null device pointer");
    const size_t need = carve_work(c, P, T, nullptr, nullptr);
    if (need > (size_t)KMC_NOALLOC_SCRATCH_BYTES)
        return set_err(KMC_E_ARG, "kmc_propose_noalloc: " + std::to_string(P) + " moving charges and " +
                                      std::to_string(T) + " proposals need " + std::to_string(need) +
                                      " bytes of scratch, more than its fixed " +
                                      std::to_string((long long)KMC_NOALLOC_SCRATCH_BYTES) + "; use kmc_propose");
    hipStream_t s = (hipStream_t)stream;
    void* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    ProposeWork w;
    carve_work(c, P, T, (unsigned char*)base, &w);
    // the bad-move counter and the tile counts start at zero (the array keeps the last
    // call's contents)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, w.err, 1, 0);
    if (w.tcount) {
        const int nt = near_tiles(c);
        hipLaunchKernelGGL(k_fill_i32, dim3((nt * TSTRIDE + 255) / 256), dim3(256), 0, s, w.tcount, nt * TSTRIDE, 0);
    }
    rc = launch_setup(c, s, P, T, d_pid, d_es, d_new_pos, d_new_q, &w);
    if (rc) return rc;
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, w.err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return set_err(KMC_E_BAD_MOVE, "Bad proposed move was detected.");
    return run_propose_core(c, s, P, T, d_dE, nullptr, precision == KMC_PREC_MIXED, 0.0, nullptr, nullptr, 0, 0, &w);
}

// exclusive scan of n int64 values over many workgroups (tmp: 2 * ceil(n / SCAN_BLOCK))
static void scan_excl_i64(hipStream_t s, int n, const int64_t* in, int64_t* out, int64_t* total, int64_t* tmp) {
    const int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    hipLaunchKernelGGL(k_scan_block_sums, dim3(nb), dim3(256), 0, s, n, in, tmp);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, s, nb, tmp, tmp + nb, total);
    hipLaunchKernelGGL(k_scan_block_apply, dim3(nb), dim3(256), 0, s, n, in, tmp + nb, out);
}

int kmc_propose_with_dats(kmc_ctx* c, void* stream, int64_t max_prop, const int64_t* d_smc, int64_t nsites,
                          const int64_t* d_sites, const double* d_ppos, const int64_t* d_masks, const double* d_pq,
                          double* d_out, int32_t precision, int64_t* h_total_moves) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (max_prop < 1) return set_err(KMC_E_ARG, "max_prop must be >= 1");
    if (nsites < 1 || !d_smc || !d_sites || !d_ppos || !d_masks || !d_out)
        return set_err(KMC_E_ARG, "bad site_max_counts / null device pointer");
    hipStream_t s = (hipStream_t)stream;
    const int n = (int)c->n;
    HIPCHK(c->d_cnt.ensure(n));
    HIPCHK(c->d_has.ensure(n));
    HIPCHK(c->d_slot.ensure(n));
    HIPCHK(c->d_off.ensure(n));
    HIPCHK(c->d_tot.ensure(2));
    HIPCHK(c->d_scan.ensure((size_t)2 * ((n + SCAN_BLOCK - 1) / SCAN_BLOCK) + 2));
    HIPCHK(hipMemsetAsync(c->err.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_dats_count, dim3((n + 255) / 256), dim3(256), 0, s, n, max_prop, d_smc, nsites, d_sites,
                       d_masks, c->d_cnt.p, c->d_has.p, c->err.p);
    scan_excl_i64(s, n, c->d_cnt.p, c->d_off.p, c->d_tot.p, c->d_scan.p);
    scan_excl_i64(s, n, c->d_has.p, c->d_slot.p, c->d_tot.p + 1, c->d_scan.p);
    int64_t tot[2] = {0, 0};
    int herr0 = 0;
    HIPCHK(hipMemcpyAsync(tot, c->d_tot.p, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&herr0, c->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr0) return fail_err_counter(c, s, KMC_E_ARG, "current_sites holds a site outside site_max_counts");
    const int64_t T = tot[0], P = tot[1];
    if (h_total_moves) *h_total_moves = T;
    if (P == 0 || T == 0) return KMC_OK;
    HIPCHK(c->d_pid.ensure(P));
    HIPCHK(c->d_es.ensure(P + 1));
    HIPCHK(c->d_npos.ensure(3 * T));
    HIPCHK(c->d_nq.ensure(T));
    HIPCHK(c->d_rate.ensure(T));
    hipLaunchKernelGGL(k_dats_scatter, dim3((n + 255) / 256), dim3(256), 0, s, n, max_prop, d_smc, d_sites, d_masks,
                       d_ppos, d_pq, c->d_cnt.p, c->d_slot.p, c->d_off.p, c->d_pid.p, c->d_es.p, c->d_npos.p,
                       c->d_nq.p, c->d_rate.p, c->d_tot.p, c->d_tot.p + 1, c->ppos4.p);
    HIPCHK(hipGetLastError());
    int rc = ensure_eval(c, T, P);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(c->err.p, 0, sizeof(int), s));
    rc = launch_setup(c, s, P, T, c->d_pid.p, c->d_es.p, c->d_npos.p, c->d_nq.p);
    if (rc) return rc;
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, c->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail_err_counter(c, s, KMC_E_BAD_MOVE, "Bad proposed move was detected.");
    return run_propose_core(c, s, P, T, d_out, c->d_rate.p, precision == KMC_PREC_MIXED);
}

int kmc_accept(kmc_ctx* c, void* stream, int64_t pid, const double* h_new_pos, double* h_energy) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (pid < 0 || pid >= c->n || !h_new_pos) return set_err(KMC_E_ARG, "bad move");
    hipStream_t s = (hipStream_t)stream;
    // host mirrors: leaf of the new position and capacity of its slots
    int mrc = refresh_mirrors(c, s);
    if (mrc) return mrc;
    const int n = c->ncps;
    const int newc = bin_axis(h_new_pos[0], c->E, c->w, n) +
                     n * (bin_axis(h_new_pos[1], c->E, c->w, n) + n * bin_axis(h_new_pos[2], c->E, c->w, n));
    const int oldc = c->h_cell[pid];
    if (newc != oldc && c->h_occ[newc] + 1 > c->occmax) {
        int rc = grow_store(c, c->occmax + std::max(4, c->occ_headroom), s);
        if (rc) return rc;
    }
    // new energy from a one-proposal FP64 propose (kmc_fmm.py:834-836)
    const int64_t es[2] = {0, 1};
    HIPCHK(hipMemcpyAsync(c->one_pid.p, &pid, sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->one_es.p, es, 2 * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->one_pos.p, h_new_pos, 3 * sizeof(double), hipMemcpyHostToDevice, s));
    int rc = kmc_propose(c, stream, 1, 1, c->one_pid.p, c->one_es.p, c->one_pos.p, nullptr, c->dE1.p, KMC_PREC_FP64);
    if (rc) return rc;
    hipLaunchKernelGGL(k_prepare_record, dim3(1), dim3(1), 0, s, c->geom(), (int)pid, h_new_pos[0], h_new_pos[1],
                       h_new_pos[2], c->ppos4.p, c->pcell.p, c->dE1.p, c->energy.p, c->rec.p);
    rc = run_accept_tail(c, s, nullptr);
    if (rc) return rc;
    int fl = 0;
    HIPCHK(hipMemcpyAsync(&fl, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (fl) {
        hipMemsetAsync(c->flags.p, 0, sizeof(int), s);
        return set_err(KMC_E_STORE, "Index was not unique, this is an error.");
    }
    c->h_occ[oldc]--;
    c->h_occ[newc]++;
    c->h_cell[pid] = newc;
    for (int d = 0; d < 3; d++) c->h_pos[3 * pid + d] = h_new_pos[d];
    if (h_energy) HIPCHK(hipMemcpy(h_energy, c->energy.p, sizeof(double), hipMemcpyDeviceToHost));
    return KMC_OK;
}

int kmc_accept_movedata(kmc_ctx* c, void* stream, const int64_t* h_movedata, double* h_energy) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (!h_movedata) return set_err(KMC_E_ARG, "null move record");
    double rd[7];
    memcpy(rd, h_movedata, sizeof(rd));
    const int64_t pid = h_movedata[7];
    if (pid < 0 || pid >= c->n) return set_err(KMC_E_ARG, "move record: charge id out of range");
    int rc = refresh_mirrors(c, (hipStream_t)stream);
    if (rc) return rc;
    const int n = c->ncps;
    const int newc = bin_axis(rd[3], c->E, c->w, n) + n * (bin_axis(rd[4], c->E, c->w, n) + n * bin_axis(rd[5], c->E, c->w, n));
    for (int d = 0; d < 3; d++)
        if (rd[d] != c->h_pos[3 * pid + d]) return set_err(KMC_E_ARG, "move record: old position is not the charge's");
    if (rd[6] != c->h_q[pid]) return set_err(KMC_E_ARG, "move record: charge is not the charge's");
    if (h_movedata[8] != c->h_cell[pid] || h_movedata[9] != newc)
        return set_err(KMC_E_ARG, "move record: cells do not match the positions");
    return kmc_accept(c, stream, pid, rd + 3, h_energy);
}

int kmc_step(kmc_ctx* c, void* stream, int64_t P, int64_t T, const int64_t* d_pid, const int64_t* d_es,
             const double* d_new_pos, double* d_dE, int64_t nsteps, const double* d_uniform, double kT,
             int64_t* d_counter, int64_t* d_selected, int32_t precision) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (P < 1 || T < 1 || T >= (int64_t)1 << 30 || nsteps < 1) return set_err(KMC_E_ARG, "bad proposal counts");
    if (!step_sized(c, P, T))
        return set_err(KMC_E_STATE, "work buffers not sized: call kmc_propose once with this T before kmc_step");
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_setup(c, s, P, T, d_pid, d_es, d_new_pos, nullptr);
    if (rc) return rc;
    rc = run_propose_core(c, s, P, T, d_dE, nullptr, precision == KMC_PREC_MIXED, kT, c->rates.p);
    if (rc) return rc;
    const SelectArgs sa = select_args(c, P, T, d_pid, d_dE, nsteps, d_uniform, kT, d_counter, d_selected);
    hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, s, sa);
    rc = run_accept_tail(c, s, d_counter);
    if (rc) return rc;
    c->mirrors_stale = true;
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

static MatArgs mat_args(kmc_ctx* c, int64_t nsteps, int64_t P, int64_t k, const int64_t* d_pids_all,
                        const double* d_disp_all, const int64_t* d_counter, int64_t* d_pid, int64_t* d_es,
                        double* d_new_pos) {
    MatArgs ma;
    ma.g = c->geom();
    ma.n = (int)c->n;
    ma.nsteps = nsteps;
    ma.P = (int)P;
    ma.k = (int)k;
    ma.pids_all = d_pids_all;
    ma.disp_all = d_disp_all;
    ma.counter = d_counter;
    ma.ppos4 = c->ppos4.p;
    ma.pid = d_pid;
    ma.es = d_es;
    ma.new_pos = d_new_pos;
    ma.flags = c->flags.p;
    return ma;
}

int kmc_materialize_moves(kmc_ctx* c, void* stream, int64_t nsteps, int64_t P, int64_t k,
                          const int64_t* d_pids_all, const double* d_disp_all, const int64_t* d_counter,
                          int64_t* d_pid, int64_t* d_es, double* d_new_pos) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first (kmc_set_particles)");
    if (P < 1 || k < 1 || nsteps < 1) return set_err(KMC_E_ARG, "bad counts");
    hipStream_t s = (hipStream_t)stream;
    const int64_t T = P * k;
    const int64_t threads = std::max(T, P + 1);
    MatArgs ma = mat_args(c, nsteps, P, k, d_pids_all, d_disp_all, d_counter, d_pid, d_es, d_new_pos);
    hipLaunchKernelGGL(k_materialize, dim3((threads + 255) / 256), dim3(256), 0, s, ma);
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

// Pipelined mode 4: the caller's stream s carries the near field, combine, select and
// the accept sweep of every step, one after the other with no wait across queues on
// that chain; side carries the accept tail and, behind it, the next step's
// materialize, set-up and long range; side2 the far field (after the set-up and the
// previous sweep, events).  Same kernels, arguments and results as one step at a time.
static int run_steps_mode4(kmc_ctx* c, hipStream_t s, int64_t nrun, int64_t P, int64_t T, const MatArgs& ma,
                           const SelectArgs& sa, int64_t* d_pid, int64_t* d_es, double* d_new_pos, double* d_dE,
                           int64_t* d_counter, double kT, bool mixed) {
    const int64_t mthreads = std::max(T, P + 1);
    const int ne = (int)(T + P);
    const bool lr = (c->bc == KMC_BC_PBC || c->bc == KMC_BC_FF_ONLY) && c->have_lr;
    const ProposeWork wk = ctx_work(c);
    HIPCHK(hipEventRecord(c->ev_fork[0], s));
    HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork[0], 0));
    int rc = KMC_OK;
    for (int64_t it = 0; it < nrun && rc == KMC_OK; it++) {
        // side: materialize, set-up, long range
        hipLaunchKernelGGL(k_materialize, dim3((mthreads + 255) / 256), dim3(256), 0, c->side, ma);
        rc = launch_setup(c, c->side, P, T, d_pid, d_es, d_new_pos, nullptr);
        if (rc) break;
        HIPCHK(hipEventRecord(c->ev_setup, c->side));
        // s: near field (after the previous sweep, stream order), recorded first of the
        // set-up's dependents
        HIPCHK(hipStreamWaitEvent(s, c->ev_setup, 0));
        launch_near_any(c, s, eval_args(c, ne, wk.epos4, wk.ecell, wk.eexcl, wk.UD), mixed, (int)P, (int)T, wk);
        // side2: far field after the set-up and (event) the previous sweep
        HIPCHK(hipStreamWaitEvent(c->side2, c->ev_setup, 0));
        if (it > 0) HIPCHK(hipStreamWaitEvent(c->side2, c->ev_sweep, 0));
        c->fns.far(eval_args(c, ne, wk.epos4, wk.ecell, wk.eexcl, wk.UI), c->side2);
        HIPCHK(hipEventRecord(c->ev_far, c->side2));
        // side: long range
        if (lr) {
            c->fns.lr(lr_args(c, P, T), c->side);
            HIPCHK(hipEventRecord(c->ev_join[0], c->side));
        }
        // s: combine, select, sweep
        HIPCHK(hipStreamWaitEvent(s, c->ev_far, 0));
        if (lr) HIPCHK(hipStreamWaitEvent(s, c->ev_join[0], 0));
        const CombineArgs ca = combine_args(c, P, T, d_dE, nullptr, kT, c->rates.p, nullptr);
        if (mixed) hipLaunchKernelGGL(k_combine<true>, dim3((T + 63) / 64), dim3(64), 0, s, ca);
        else hipLaunchKernelGGL(k_combine<false>, dim3((T + 63) / 64), dim3(64), 0, s, ca);
        hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, s, sa);
        HIPCHK(hipEventRecord(c->ev_fork[1], s));
        c->fns.sweep(sweep_rec_args(c), (int)c->ncells, s, c->sweep_block, c->sweep_pairs, c->sweep_split);
        HIPCHK(hipEventRecord(c->ev_sweep, s));
        // side: the accept tail after the select (beside the sweep)
        HIPCHK(hipStreamWaitEvent(c->side, c->ev_fork[1], 0));
        c->fns.md(tail_args(c, d_counter), c->side);
        HIPCHK(hipGetLastError());
    }
    // join side (its last tail, or the work a failed step left) and side2 into s
    hipError_t e = hipEventRecord(c->ev_join[1], c->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, c->ev_join[1], 0);
    if (e == hipSuccess) e = hipEventRecord(c->ev_fork2, c->side2);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, c->ev_fork2, 0);
    if (rc == KMC_OK && e != hipSuccess) return set_err(KMC_E_HIP, std::string("join: ") + hipGetErrorString(e));
    return rc;
}

int kmc_run_steps(kmc_ctx* c, void* stream, int64_t nrun, int64_t nsteps_moves, int64_t P, int64_t k,
                  const int64_t* d_pids_all, const double* d_disp_all, int64_t nsteps_uniform,
                  const double* d_uniform, double kT, int64_t* d_counter, int64_t* d_selected, int64_t* d_pid,
                  int64_t* d_es, double* d_new_pos, double* d_dE, int32_t precision, int32_t pipelined) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (nrun < 0 || P < 1 || k < 1 || nsteps_moves < 1 || nsteps_uniform < 1 || P * k >= (int64_t)1 << 30)
        return set_err(KMC_E_ARG, "bad counts");
    const int64_t T = P * k;
    if (!step_sized(c, P, T))
        return set_err(KMC_E_STATE, "work buffers not sized: call kmc_propose once with this T before kmc_run_steps");
    if (pipelined < 0 || pipelined > 4) return set_err(KMC_E_ARG, "pipelined must be 0, 1, 2, 3 or 4");
    hipStream_t s = (hipStream_t)stream;
    const bool pipe = pipelined != 0 && c->bc != KMC_BC_FF_ONLY;
    const bool mixed = precision == KMC_PREC_MIXED;
    const int64_t mthreads = std::max(T, P + 1);
    const MatArgs ma = mat_args(c, nsteps_moves, P, k, d_pids_all, d_disp_all, d_counter, d_pid, d_es, d_new_pos);
    const SelectArgs sa = select_args(c, P, T, d_pid, d_dE, nsteps_uniform, d_uniform, kT, d_counter, d_selected);
    if (pipe && pipelined == 4) {
        const int rc4 = run_steps_mode4(c, s, nrun, P, T, ma, sa, d_pid, d_es, d_new_pos, d_dE, d_counter, kT, mixed);
        if (rc4) return rc4;
        c->mirrors_stale = true;
        HIPCHK(hipGetLastError());
        return KMC_OK;
    }
    int rc = KMC_OK;
    int64_t done = 0;   // steps whose sweep may be pending on side2
    for (int64_t it = 0; it < nrun && rc == KMC_OK; it++) {
        hipLaunchKernelGGL(k_materialize, dim3((mthreads + 255) / 256), dim3(256), 0, s, ma);
        rc = launch_setup(c, s, P, T, d_pid, d_es, d_new_pos, nullptr);
        if (rc == KMC_OK)
            rc = run_propose_core(c, s, P, T, d_dE, nullptr, mixed, kT, c->rates.p, nullptr, 0,
                                  (pipe && it > 0) ? (pipelined >= 2 ? 2 : 1) : 0, nullptr, pipe,
                                  pipe && pipelined == 3);
        if (rc == KMC_OK) {
            hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, s, sa);
            rc = run_accept_tail(c, s, d_counter, pipe);
            if (rc == KMC_OK) done = it + 1;
        }
    }
    // join the last deferred sweep into the call's stream, also when a step failed (a
    // capture must end joined, and later work on s must follow the sweep); mode 3 also
    // the last far field (side2's work follows the sweeps there)
    if (pipe && done > 0) {
        hipError_t e = hipStreamWaitEvent(s, c->ev_sweep, 0);
        if (e == hipSuccess && pipelined == 3) e = hipStreamWaitEvent(s, c->ev_far, 0);
        if (rc == KMC_OK && e != hipSuccess)
            return set_err(KMC_E_HIP, std::string("hipStreamWaitEvent: ") + hipGetErrorString(e));
    }
    if (rc) return rc;
    c->mirrors_stale = true;
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

int kmc_check_flags(kmc_ctx* c, void* stream, int32_t* h_flags) {
    if (check_ctx(c)) return KMC_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int fl = 0;
    HIPCHK(hipMemcpyAsync(&fl, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemset(c->flags.p, 0, sizeof(int)));
    if (h_flags) *h_flags = fl;
    return KMC_OK;
}

// room for device-side steps: every leaf gets at least `slots` free slots (the cell
// tables grow, keeping their contents, to the largest occupancy + slots + the context's
// headroom when the fullest leaf has fewer); synchronises the stream
int kmc_store_reserve(kmc_ctx* c, void* stream, int32_t slots, int32_t* h_grew) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (slots < 0 || slots > (1 << 20)) return set_err(KMC_E_ARG, "slots must be in [0, 2^20]");
    if (h_grew) *h_grew = 0;
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(c->occpeak.ensure(1));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, c->occpeak.p, 1, 0);
    const int nb = (int)std::min<int64_t>(1024, (c->ncells + 255) / 256);
    hipLaunchKernelGGL(k_occ_peak, dim3(nb), dim3(256), 0, s, c->occ.p, (int)c->ncells, c->occpeak.p);
    HIPCHK(hipGetLastError());
    int peak = 0;
    HIPCHK(hipMemcpyAsync(&peak, c->occpeak.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (c->occmax - peak >= slots) return KMC_OK;
    const int rc = grow_store(c, peak + slots + c->occ_headroom, s);
    if (rc) return rc;
    if (h_grew) *h_grew = 1;
    return KMC_OK;
}

int kmc_time_kernel(kmc_ctx* c, void* stream, int32_t which, int32_t reps, int32_t precision, double* h_ms) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (reps < 1 || !h_ms) return set_err(KMC_E_ARG, "bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int64_t P = c->last_P, T = c->last_T;
    if ((which != 0 && (P < 1 || T < 1)) || !c->initialised) return set_err(KMC_E_STATE, "run a propose first");
    const bool mixed = precision == KMC_PREC_MIXED;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++) {
        if (which == 0) {
            SweepArgs sa;
            sa.g = c->geom();
            sa.lexp = c->lexp.p;
            sa.ncells = (int)c->ncells;
            sa.mode = 0;
            sa.rec = c->rec.p;
            sa.src_pos4 = nullptr;
            sa.src_cell = nullptr;
            sa.src_begin = sa.src_end = 0;
            c->fns.sweep(sa, (int)c->ncells, s, c->sweep_block, c->sweep_pairs, c->sweep_split);
        } else if (which == 1) {
            EvalArgs an = eval_args(c, (int)(T + P), c->epos4.p, c->ecell.p, c->eexcl.p, c->UD.p);
            launch_near_any(c, s, an, mixed, (int)P, (int)T, ctx_work(c));
        } else if (which == 2) {
            EvalArgs af = eval_args(c, (int)(T + P), c->epos4.p, c->ecell.p, c->eexcl.p, c->UI.p);
            c->fns.far(af, s);
        } else if (which == 3) {
            if (c->have_lr && (size_t)12 * P * lr_split(P, nullptr) <= c->cold.n) c->fns.lr(lr_args(c, P, T), s);
        } else {
            return set_err(KMC_E_ARG, "unknown kernel id");
        }
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *h_ms = (double)ms / reps;
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

// roofline denominators measured on the device: a streaming copy (HBM read + write)
// and a loop of independent FP64 FMA chains (8 per lane, 2 flops each)
__global__ void __launch_bounds__(256) k_probe_copy(const double4* __restrict__ src, double4* __restrict__ dst,
                                                     int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}
constexpr int PROBE_CHAINS = 8;
__global__ void __launch_bounds__(256) k_probe_fma(double* __restrict__ out, int iters, double m, double c) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    double a[PROBE_CHAINS];
#pragma unroll
    for (int j = 0; j < PROBE_CHAINS; j++) a[j] = 1e-3 * (double)(gid + j);
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < PROBE_CHAINS; j++) a[j] = fma(a[j], m, c);
    }
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < PROBE_CHAINS; j++) t += a[j];
    out[gid] = t;
}

int kmc_probe_peaks(void* stream, int64_t bytes, int32_t reps, double* h_out) {
    if (bytes < (1 << 20) || bytes > ((int64_t)1 << 34) || reps < 1 || !h_out)
        return set_err(KMC_E_ARG, "bytes must be in [1 MiB, 16 GiB], reps >= 1");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n4 = bytes / (int64_t)sizeof(double4);
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int copy_blocks = ncu * 16;
    const int fma_blocks = ncu * 8, fma_iters = 4096;
    DBuf<double4> src, dst;
    DBuf<double> out;
    HIPCHK(src.ensure(n4));
    HIPCHK(dst.ensure(n4));
    HIPCHK(out.ensure((size_t)fma_blocks * 256));
    HIPCHK(hipMemsetAsync(src.p, 0, (size_t)n4 * sizeof(double4), s));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float ms = 0.0f;
    // warm-up, then timed
    hipLaunchKernelGGL(k_probe_copy, dim3(copy_blocks), dim3(256), 0, s, src.p, dst.p, n4);
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++)
        hipLaunchKernelGGL(k_probe_copy, dim3(copy_blocks), dim3(256), 0, s, src.p, dst.p, n4);
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    h_out[0] = 2.0 * (double)n4 * sizeof(double4) * reps / (ms * 1e-3) / 1e9;
    hipLaunchKernelGGL(k_probe_fma, dim3(fma_blocks), dim3(256), 0, s, out.p, fma_iters, 0.999999, 1e-7);
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++)
        hipLaunchKernelGGL(k_probe_fma, dim3(fma_blocks), dim3(256), 0, s, out.p, fma_iters, 0.999999, 1e-7);
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    const double flops = 2.0 * PROBE_CHAINS * (double)fma_iters * fma_blocks * 256.0 * reps;
    h_out[1] = flops / (ms * 1e-3) / 1e12;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

int kmc_eval_field(kmc_ctx* c, void* stream, int64_t n, const double* d_points, double* d_out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (n < 0 || n >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad point count");
    if (n == 0) return KMC_OK;
    if (!d_points || !d_out) return set_err(KMC_E_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    const int ni = (int)n, nfar = far_parts(c->L);
    DBuf<double4> pos4;
    DBuf<int32_t> cell, excl;
    DBuf<double> ud, ui, lr;
    HIPCHK(pos4.ensure(ni));
    HIPCHK(cell.ensure(ni));
    HIPCHK(excl.ensure(ni));
    hipLaunchKernelGGL(k_eval_points, dim3((ni + 255) / 256), dim3(256), 0, s, c->geom(), ni, d_points, pos4.p, cell.p,
                       excl.p);
    const bool near_far = c->bc != KMC_BC_FF_ONLY;
    const bool lrt = (c->bc == KMC_BC_PBC || c->bc == KMC_BC_FF_ONLY) && c->have_lr;
    if (near_far) {
        HIPCHK(ud.ensure(ni));
        HIPCHK(ui.ensure((size_t)ni * nfar));
        launch_near(eval_args(c, ni, pos4.p, cell.p, excl.p, ud.p), c->near_lanes, false, s);
        c->fns.far(eval_args(c, ni, pos4.p, cell.p, excl.p, ui.p), s);
    }
    if (lrt) {
        HIPCHK(lr.ensure(ni));
        hipLaunchKernelGGL(k_eval_lr, dim3((ni + 255) / 256), dim3(256), 0, s, c->L, ni, pos4.p, c->gvec.p, lr.p);
    }
    hipLaunchKernelGGL(k_eval_combine, dim3((ni + 255) / 256), dim3(256), 0, s, ni, near_far ? ud.p : nullptr,
                       near_far ? ui.p : nullptr, nfar, lrt ? lr.p : nullptr, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return KMC_OK;
}

int kmc_sweep_hop(kmc_ctx* c, void* stream, double q, const double* h_old, const double* h_new) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first (kmc_set_particles)");
    if (c->bc == KMC_BC_FF_ONLY) return KMC_OK;
    hipStream_t s = (hipStream_t)stream;
    MoveRec r{};
    const int n = c->ncps;
    for (int d = 0; d < 3; d++) {
        r.old_pos[d] = h_old[d];
        r.new_pos[d] = h_new[d];
    }
    r.q = q;
    r.old_cell = bin_axis(h_old[0], c->E, c->w, n) + n * (bin_axis(h_old[1], c->E, c->w, n) + n * bin_axis(h_old[2], c->E, c->w, n));
    r.new_cell = bin_axis(h_new[0], c->E, c->w, n) + n * (bin_axis(h_new[1], c->E, c->w, n) + n * bin_axis(h_new[2], c->E, c->w, n));
    r.valid = 1;
    HIPCHK(hipMemcpyAsync(c->rec.p, &r, sizeof(MoveRec), hipMemcpyHostToDevice, s));
    SweepArgs sa;
    sa.g = c->geom();
    sa.lexp = c->lexp.p;
    sa.ncells = (int)c->ncells;
    sa.mode = 0;
    sa.rec = c->rec.p;
    sa.src_pos4 = nullptr;
    sa.src_cell = nullptr;
    sa.src_begin = sa.src_end = 0;
    c->fns.sweep(sa, (int)c->ncells, s, c->sweep_block, c->sweep_pairs, c->sweep_split);
    HIPCHK(hipGetLastError());
    return KMC_OK;
}


int kmc_set_particle_count(kmc_ctx* c, int64_t n) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (n < 1 || n >= (int64_t)1 << 31) return set_err(KMC_E_ARG, "n out of range");
    c->n = n;
    c->have_particles = false;
    c->initialised = false;
    c->generation++;
    return KMC_OK;
}

int kmc_particle_count(kmc_ctx* c, int64_t* out) {
    if (check_ctx(c) || !out) return set_err(KMC_E_ARG, "null argument");
    *out = c->n;
    return KMC_OK;
}

// charges' energies (energy_unit applied) or, potential: the potential at each charge
// from all the others (unit charge, no energy_unit)
static int charge_entries(kmc_ctx* c, void* stream, int64_t n, const int64_t* d_ids, double* d_out, bool potential) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (n < 0 || n >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad charge count");
    if (n == 0) return KMC_OK;
    if (!d_ids || !d_out) return set_err(KMC_E_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    const int ni = (int)n;
    DBuf<double4> pos4;
    DBuf<int32_t> cell, excl;
    DBuf<double> ud, ui, lr;
    HIPCHK(pos4.ensure(ni));
    HIPCHK(cell.ensure(ni));
    HIPCHK(excl.ensure(ni));
    HIPCHK(hipMemsetAsync(c->err.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_gather_ids, dim3((ni + 255) / 256), dim3(256), 0, s, c->geom(), ni, d_ids, c->n, c->ppos4.p,
                       c->pcell.p, 1.0, pos4.p, cell.p, excl.p, c->err.p);
    if (potential) hipLaunchKernelGGL(k_unit_charges, dim3((ni + 255) / 256), dim3(256), 0, s, ni, pos4.p);
    int rc = entry_energies(c, s, ni, pos4.p, cell.p, excl.p, d_out, ud, ui, lr, potential ? 1.0 : c->eu);
    if (rc) return rc;
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, c->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail_err_counter(c, s, KMC_E_ARG, "charge id out of range");
    return KMC_OK;
}

int kmc_charge_energies(kmc_ctx* c, void* stream, int64_t n, const int64_t* d_ids, double* d_out) {
    return charge_entries(c, stream, n, d_ids, d_out, false);
}

int kmc_charge_potentials(kmc_ctx* c, void* stream, int64_t n, const int64_t* d_ids, double* d_out) {
    return charge_entries(c, stream, n, d_ids, d_out, true);
}

int kmc_group_energy(kmc_ctx* c, void* stream, int64_t ngroups, int64_t m, const double* d_pos, const double* d_q,
                     double* d_out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->have_particles) return set_err(KMC_E_STATE, "set particles first (kmc_set_particles)");
    if ((c->bc == KMC_BC_PBC || c->bc == KMC_BC_FF_ONLY) && !c->have_lr)
        return set_err(KMC_E_STATE, "set the long-range operator first (pbc / ff-only)");
    if (ngroups < 0 || m < 0 || ngroups * m >= (int64_t)1 << 30 || ngroups >= (int64_t)1 << 31)
        return set_err(KMC_E_ARG, "bad group sizes");
    if (ngroups == 0) return KMC_OK;
    if (!d_pos || !d_q || !d_out) return set_err(KMC_E_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    const int n = (int)(ngroups * m);
    DBuf<double4> pos4;
    DBuf<int32_t> cell;
    HIPCHK(pos4.ensure(std::max(n, 1)));
    HIPCHK(cell.ensure(std::max(n, 1)));
    if (n > 0)
        hipLaunchKernelGGL(k_pack_charges, dim3((n + 255) / 256), dim3(256), 0, s, c->geom(), n, d_pos, d_q, pos4.p,
                           cell.p, nullptr);
    launch_group_energy(c, s, (int)ngroups, (int)m, pos4.p, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return KMC_OK;
}

int kmc_propose_extract(kmc_ctx* c, void* stream, int64_t nrows, int64_t m, const int64_t* d_ids,
                        int32_t self_energy, double* d_out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (nrows < 0 || m < 0 || nrows * m >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad id array shape");
    if (nrows == 0) return KMC_OK;
    if (!d_out || (m > 0 && !d_ids)) return set_err(KMC_E_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    if (m == 0) {
        hipLaunchKernelGGL(k_extract_rows, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, s, (int)nrows, 0,
                           nullptr, nullptr, d_out);
        HIPCHK(hipStreamSynchronize(s));
        return KMC_OK;
    }
    return propose_extract_core(c, s, (int)nrows, (int)m, d_ids, self_energy != 0, d_out);
}

int kmc_propose_inject(kmc_ctx* c, void* stream, int64_t n, const double* d_pos, const double* d_q, double* d_out) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (n < 1 || n >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad charge count");
    if (!d_pos || !d_q || !d_out) return set_err(KMC_E_ARG, "null pointer");
    return propose_inject_core(c, (hipStream_t)stream, (int)n, d_pos, d_q, d_out);
}

int kmc_extract(kmc_ctx* c, void* stream, int64_t k, const int64_t* h_ids, double* h_energy) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (k < 0) return set_err(KMC_E_ARG, "bad charge count");
    hipStream_t s = (hipStream_t)stream;
    if (k > 0) {
        if (!h_ids) return set_err(KMC_E_ARG, "null pointer");
        if (k >= c->n) return set_err(KMC_E_ARG, "cannot extract every charge of the system");
        std::vector<int64_t> desc(h_ids, h_ids + k);
        std::sort(desc.begin(), desc.end(), [](int64_t a, int64_t b) { return a > b; });
        for (int64_t t = 0; t < k; t++) {
            if (desc[t] < 0 || desc[t] >= c->n) return set_err(KMC_E_ARG, "charge id out of range");
            if (t > 0 && desc[t] == desc[t - 1]) return set_err(KMC_E_ARG, "charge ids to extract must be unique");
        }
        int rc = refresh_mirrors(c, s);
        if (rc) return rc;
        const int n0 = (int)c->n, kk = (int)k;
        DBuf<int64_t> d_ids, d_desc;
        DBuf<double> dE;
        DBuf<double4> src;
        DBuf<int32_t> scell, sexcl;
        HIPCHK(d_ids.ensure(kk));
        HIPCHK(d_desc.ensure(kk));
        HIPCHK(dE.ensure(1));
        HIPCHK(src.ensure(kk));
        HIPCHK(scell.ensure(kk));
        HIPCHK(sexcl.ensure(kk));
        HIPCHK(hipMemcpyAsync(d_ids.p, h_ids, kk * sizeof(int64_t), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_desc.p, desc.data(), kk * sizeof(int64_t), hipMemcpyHostToDevice, s));
        // energy change of removing the charges as one group (kmc_inject_extract.py:240-251)
        rc = propose_extract_core(c, s, 1, kk, d_ids.p, true, dE.p);
        if (rc) return rc;
        hipLaunchKernelGGL(k_add_energy, dim3(1), dim3(1), 0, s, dE.p, c->energy.p);
        // the charges with negated sign as sources: leaf expansions
        // (accept_local_exp in extraction mode, kmc_octal.py:1091-1217) and level-0 multipole
        hipLaunchKernelGGL(k_gather_ids, dim3((kk + 255) / 256), dim3(256), 0, s, c->geom(), kk, d_ids.p, c->n,
                           c->ppos4.p, c->pcell.p, -1.0, src.p, scell.p, sexcl.p, c->err.p);
        if (c->bc != KMC_BC_FF_ONLY) sweep_sources(c, s, src.p, scell.p, kk);
        if (uses_lr(c)) {
            rc = lr_add_sources(c, s, src.p, kk);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_extract_store, dim3(1), dim3(64), 0, s, kk, d_desc.p, n0, c->occ.p, c->slot_pos.p,
                           c->slot_id.p, c->occmax, c->ppos4.p, c->pcell.p, c->flags.p);
        HIPCHK(hipGetLastError());
        int fl = 0;
        HIPCHK(hipMemcpyAsync(&fl, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (fl) {
            hipMemsetAsync(c->flags.p, 0, sizeof(int), s);
            return set_err(KMC_E_STORE, "Index was not unique, this is an error.");
        }
        // host mirrors, the same compaction
        int n = n0;
        for (int t = 0; t < kk; t++) {
            const int id = (int)desc[t];
            c->h_occ[c->h_cell[id]]--;
            const int last = n - 1;
            if (id != last) {
                for (int d = 0; d < 3; d++) c->h_pos[3 * id + d] = c->h_pos[3 * last + d];
                c->h_q[id] = c->h_q[last];
                c->h_cell[id] = c->h_cell[last];
            }
            n = last;
        }
        c->h_pos.resize(3 * (size_t)n);
        c->h_q.resize(n);
        c->h_cell.resize(n);
        c->n = n;
        c->generation++;
    }
    if (h_energy) HIPCHK(hipMemcpy(h_energy, c->energy.p, sizeof(double), hipMemcpyDeviceToHost));
    return KMC_OK;
}

int kmc_inject(kmc_ctx* c, void* stream, int64_t k, const double* h_pos, const double* h_q, double* h_energy) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    if (k < 0 || c->n + k >= (int64_t)1 << 31) return set_err(KMC_E_ARG, "bad charge count");
    hipStream_t s = (hipStream_t)stream;
    if (k > 0) {
        if (!h_pos || !h_q) return set_err(KMC_E_ARG, "null pointer");
        const double half = 0.5 * c->E;
        for (int64_t i = 0; i < 3 * k; i++)
            if (!(h_pos[i] >= -half && h_pos[i] <= half))
                return set_err(KMC_E_BAD_MOVE, "Injected position is outside simulation domain.");
        int rc = refresh_mirrors(c, s);
        if (rc) return rc;
        const int n0 = (int)c->n, kk = (int)k;
        // leaves of the new charges and room in their cells
        std::vector<int32_t> cells(kk);
        std::unordered_map<int32_t, int> add;
        int need = 0;
        for (int i = 0; i < kk; i++) {
            const int nc = c->ncps;
            cells[i] = bin_axis(h_pos[3 * i], c->E, c->w, nc) +
                       nc * (bin_axis(h_pos[3 * i + 1], c->E, c->w, nc) + nc * bin_axis(h_pos[3 * i + 2], c->E, c->w, nc));
            need = std::max(need, c->h_occ[cells[i]] + ++add[cells[i]]);
        }
        DBuf<double> d_pos, d_q, dE;
        DBuf<double4> src;
        DBuf<int32_t> scell;
        HIPCHK(d_pos.ensure(3 * (size_t)kk));
        HIPCHK(d_q.ensure(kk));
        HIPCHK(dE.ensure(1));
        HIPCHK(src.ensure(kk));
        HIPCHK(scell.ensure(kk));
        HIPCHK(hipMemcpyAsync(d_pos.p, h_pos, 3 * kk * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_q.p, h_q, kk * sizeof(double), hipMemcpyHostToDevice, s));
        // energy change (kmc_inject_extract.py:304-349)
        rc = propose_inject_core(c, s, kk, d_pos.p, d_q.p, dE.p);
        if (rc) return rc;
        hipLaunchKernelGGL(k_add_energy, dim3(1), dim3(1), 0, s, dE.p, c->energy.p);
        if (need > c->occmax) {
            rc = grow_store(c, need + std::max(4, c->occ_headroom), s);
            if (rc) return rc;
        }
        rc = grow_keep(c, c->ppos4, n0, (size_t)n0 + kk, s);
        if (rc) return rc;
        rc = grow_keep(c, c->pcell, n0, (size_t)n0 + kk, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_pack_charges, dim3((kk + 255) / 256), dim3(256), 0, s, c->geom(), kk, d_pos.p, d_q.p, src.p,
                           scell.p, nullptr);
        hipLaunchKernelGGL(k_inject_store, dim3(1), dim3(64), 0, s, kk, n0, src.p, scell.p, c->occ.p, c->slot_pos.p,
                           c->slot_id.p, c->occmax, c->ppos4.p, c->pcell.p, c->flags.p);
        // leaf expansions (accept_local_exp in injection mode, kmc_octal.py:1091-1217) and
        // level-0 multipole
        if (c->bc != KMC_BC_FF_ONLY) sweep_sources(c, s, src.p, scell.p, kk);
        if (uses_lr(c)) {
            rc = lr_add_sources(c, s, src.p, kk);
            if (rc) return rc;
        }
        HIPCHK(hipGetLastError());
        int fl = 0;
        HIPCHK(hipMemcpyAsync(&fl, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (fl) {
            hipMemsetAsync(c->flags.p, 0, sizeof(int), s);
            return set_err(KMC_E_STORE, "cell store overflow while injecting");
        }
        for (int i = 0; i < kk; i++) {
            for (int d = 0; d < 3; d++) c->h_pos.push_back(h_pos[3 * i + d]);
            c->h_q.push_back(h_q[i]);
            c->h_cell.push_back(cells[i]);
            c->h_occ[cells[i]]++;
        }
        c->n = n0 + kk;
        c->generation++;
    }
    if (h_energy) HIPCHK(hipMemcpy(h_energy, c->energy.p, sizeof(double), hipMemcpyDeviceToHost));
    return KMC_OK;
}

int kmc_lattice_step(kmc_ctx* c, void* stream, int64_t M, const double* h_offsets, double tol2, int32_t wrap,
                     double* d_new_pos, int64_t* d_masks, double* d_dE, int64_t nsteps, const double* d_uniform,
                     double kT, int32_t rate_rule, int64_t* d_counter, int64_t* d_selected, int32_t precision) {
    if (check_ctx(c)) return KMC_E_ARG;
    if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before this call");
    const int64_t P = c->n, T = c->n * M;
    if (M < 1 || M > 4096 || T >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad number of offsets");
    if (!h_offsets || !d_new_pos || !d_masks || !d_dE || !d_uniform || !d_counter || !d_selected)
        return set_err(KMC_E_ARG, "null pointer");
    if (!(tol2 >= 0.0) || tol2 > c->w * c->w) return set_err(KMC_E_ARG, "overlap tolerance must lie in [0, leaf width^2]");
    if (nsteps < 1 || !(kT > 0.0) || (rate_rule != 0 && rate_rule != 1))
        return set_err(KMC_E_ARG, "bad nsteps, kT or rate rule");
    for (int64_t k = 0; k < 3 * M; k++)
        if (!(fabs(h_offsets[k]) <= c->max_move))
            return set_err(KMC_E_BAD_MOVE, "offset component beyond max_move");
    hipStream_t s = (hipStream_t)stream;
    const int n = (int)P, nt = (int)T;
    // sizing (allocates: the first call for a charge count and offset set runs eagerly,
    // later calls may be captured)
    const bool same_off = c->h_lat_off.size() == (size_t)(3 * M) &&
                          std::equal(c->h_lat_off.begin(), c->h_lat_off.end(), h_offsets);
    if (!step_sized(c, P, T) || c->lat_n != P || !same_off) {
        int rc = ensure_eval(c, T, P);
        if (rc) return rc;
        if (uses_lr(c)) {
            rc = ensure_lr_work(c, P);
            if (rc) return rc;
        }
        HIPCHK(c->lat_pid.ensure(P));
        HIPCHK(c->lat_es.ensure(P + 1));
        HIPCHK(c->lat_off.ensure(3 * M));
        hipLaunchKernelGGL(k_lattice_index, dim3((n + 256) / 256), dim3(256), 0, s, n, (int)M, c->lat_pid.p,
                           c->lat_es.p);
        HIPCHK(hipMemcpyAsync(c->lat_off.p, h_offsets, 3 * M * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        c->h_lat_off.assign(h_offsets, h_offsets + 3 * M);
        c->lat_n = P;
    }
    hipLaunchKernelGGL(k_lattice_proposals, dim3((nt + 255) / 256), dim3(256), 0, s, n, (int)M, c->E, wrap,
                       c->ppos4.p, c->lat_off.p, d_new_pos, d_masks);
    hipLaunchKernelGGL(k_overlap_mask, dim3((nt + 255) / 256), dim3(256), 0, s, c->geom(), n, (int)M, tol2, wrap,
                       d_new_pos, c->occ.p, c->slot_pos.p, c->slot_id.p, c->occmax, d_masks);
    // masked proposals (outside the domain without wrapping, or overlapping) are not moves
    int rc = launch_setup(c, s, P, T, c->lat_pid.p, c->lat_es.p, d_new_pos, nullptr, nullptr, d_masks);
    if (rc) return rc;
    rc = run_propose_core(c, s, P, T, d_dE, nullptr, precision == KMC_PREC_MIXED, kT, c->rates.p, d_masks, rate_rule);
    if (rc) return rc;
    const SelectArgs sa = select_args(c, P, T, c->lat_pid.p, d_dE, nsteps, d_uniform, kT, d_counter, d_selected);
    hipLaunchKernelGGL(k_select, dim3(1), dim3(1024), 0, s, sa);
    rc = run_accept_tail(c, s, d_counter);
    if (rc) return rc;
    c->mirrors_stale = true;
    HIPCHK(hipGetLastError());
    return KMC_OK;
}

int kmc_discover_sites(void* stream, int64_t n, const double* d_pos, int64_t n_inject, const double* d_inject_sites,
                       int64_t n_extract, const double* d_extract_sites, double tol2, int64_t* d_inject_flags,
                       int64_t* d_extract_flags) {
    if (n < 0 || n_inject < 0 || n_extract < 0 || n_inject >= (int64_t)1 << 30 || n_extract >= (int64_t)1 << 30)
        return set_err(KMC_E_ARG, "bad sizes");
    if ((n > 0 && (!d_pos || !d_extract_flags)) || (n_inject > 0 && (!d_inject_sites || !d_inject_flags)) ||
        (n_extract > 0 && !d_extract_sites))
        return set_err(KMC_E_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    if (n_inject > 0)
        hipLaunchKernelGGL(k_zero_i64, dim3((unsigned)((n_inject + 255) / 256)), dim3(256), 0, s, n_inject,
                           d_inject_flags);
    if (n > 0)
        hipLaunchKernelGGL(k_discover_sites, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d_pos,
                           (int)n_inject, d_inject_sites, (int)n_extract, d_extract_sites, tol2,
                           (unsigned long long*)d_inject_flags, d_extract_flags);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return KMC_OK;
}

}  // extern "C"

// ============================================================================
// batches of independent systems: one launch per kernel kind (and order) for all
// ============================================================================
struct kmc_batch {
    struct Sys {
        kmc_ctx* c = nullptr;
        int64_t nsteps = 0, P = 0, k = 0;
        const int64_t* pids_all = nullptr;
        const double* disp_all = nullptr;
        const double* uniform = nullptr;
        int64_t* counter = nullptr;
        int64_t* selected = nullptr;
        int64_t generation = -1;       // context layout the argument tables were built on
        int64_t* pid = nullptr;
        int64_t* es = nullptr;
        double* new_pos = nullptr;
        double* dE = nullptr;
        bool set = false;
    };
    struct Group {                    // systems of one order L and one periodicity (ff-only apart)
        int L = 0;
        bool per = false;
        bool ff = false;              // ff-only systems: long range only (no near / far field, no sweep)
        OrderFns fns{};
        BatchShape shape{}, lr_shape{};
        DBuf<EvalArgs> far;
        DBuf<LR2Args> lr;
        DBuf<SweepArgs> sweep;
        DBuf<TailArgs> tail;
        // blocked sweep (every system of the group set to sweep_block 1): flat table of
        // (system, block) pairs
        bool blocked = false;
        DBuf<int2> sweep_tab;
        int sweep_nflat = 0;
    };
    std::vector<Sys> sys;
    std::vector<Group*> groups;
    double kT = 1.0;
    bool mixed = true;
    bool ready = false;
    int near_lanes = 27;  // the systems' common per-point lane count (1..64), else 27: k_near_cells' layout
    BatchShape all{};
    DBuf<MatArgs> mat;
    DBuf<SetupArgs> setup;
    DBuf<EvalArgs> nearA;
    DBuf<CombineArgs> comb;
    DBuf<SelectArgs> sel;
    // side streams for the per-group launches (far field, long range, sweep + tail of
    // one (order, periodicity) group each), forked from and joined back to the step's
    // stream with events so that a captured step keeps the dependencies
    std::vector<hipStream_t> side;
    hipEvent_t ev_fork[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_join;   // two per side stream (far/long-range phase, sweep phase)
    ~kmc_batch() {
        for (Group* g : groups) delete g;
        for (hipStream_t x : side) hipStreamDestroy(x);
        for (hipEvent_t e : ev_join) hipEventDestroy(e);
        for (hipEvent_t e : ev_fork)
            if (e) hipEventDestroy(e);
    }
};

int kmc_batch_create(kmc_batch** out, kmc_ctx* const* ctxs, int32_t n) {
    if (!out) return set_err(KMC_E_ARG, "null output pointer");
    *out = nullptr;
    if (n < 1 || n > 65535 || !ctxs) return set_err(KMC_E_ARG, "batch size must be 1 .. 65535");
    for (int i = 0; i < n; i++) {
        kmc_ctx* c = ctxs[i];
        if (!c) return set_err(KMC_E_ARG, "null context in batch");
        if (!c->initialised) return set_err(KMC_E_STATE, "Run initialise before batching a context");
        if (c->ncells % 64) return set_err(KMC_E_UNSUPPORTED, "batched systems need at least 4^3 leaves (r >= 3)");
    }
    kmc_batch* b = new kmc_batch();
    b->sys.resize(n);
    for (int i = 0; i < n; i++) b->sys[i].c = ctxs[i];
    *out = b;
    return KMC_OK;
}

int kmc_batch_destroy(kmc_batch* b) {
    delete b;
    return KMC_OK;
}

int kmc_batch_set_streams(kmc_batch* b, int32_t nstreams) {
    if (!b) return set_err(KMC_E_ARG, "null batch");
    if (nstreams < 1 || nstreams > 16) return set_err(KMC_E_ARG, "nstreams must be in [1, 16]");
    while ((int)b->side.size() > nstreams - 1) {
        HIPCHK(hipStreamDestroy(b->side.back()));
        b->side.pop_back();
        for (int k = 0; k < 2; k++) {
            HIPCHK(hipEventDestroy(b->ev_join.back()));
            b->ev_join.pop_back();
        }
    }
    for (int i = 0; i < 2; i++)
        if (!b->ev_fork[i]) HIPCHK(hipEventCreateWithFlags(&b->ev_fork[i], hipEventDisableTiming));
    while ((int)b->side.size() < nstreams - 1) {
        hipStream_t x = nullptr;
        hipEvent_t e = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        b->side.push_back(x);
        for (int k = 0; k < 2; k++) {   // one join event per side stream and phase
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->ev_join.push_back(e);
        }
    }
    return KMC_OK;
}

int kmc_batch_set_system(kmc_batch* b, int32_t i, int64_t nsteps, int64_t P, int64_t k, const int64_t* d_pids_all,
                         const double* d_disp_all, const double* d_uniform, int64_t* d_counter, int64_t* d_selected,
                         int64_t* d_pid, int64_t* d_es, double* d_new_pos, double* d_dE) {
    if (!b) return set_err(KMC_E_ARG, "null batch");
    if (i < 0 || i >= (int)b->sys.size()) return set_err(KMC_E_ARG, "system index out of range");
    if (nsteps < 1 || P < 1 || k < 1 || P * k >= (int64_t)1 << 30) return set_err(KMC_E_ARG, "bad counts");
    if (!d_pids_all || !d_disp_all || !d_uniform || !d_counter || !d_selected || !d_pid || !d_es || !d_new_pos || !d_dE)
        return set_err(KMC_E_ARG, "null device pointer");
    kmc_batch::Sys& y = b->sys[i];
    y.nsteps = nsteps;
    y.P = P;
    y.k = k;
    y.pids_all = d_pids_all;
    y.disp_all = d_disp_all;
    y.uniform = d_uniform;
    y.counter = d_counter;
    y.selected = d_selected;
    y.pid = d_pid;
    y.es = d_es;
    y.new_pos = d_new_pos;
    y.dE = d_dE;
    y.set = true;
    b->ready = false;
    return KMC_OK;
}

template <typename T>
static hipError_t upload(DBuf<T>& d, const std::vector<T>& h) {
    hipError_t e = d.ensure(std::max<size_t>(h.size(), 1));
    if (e != hipSuccess) return e;
    return hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
}

int kmc_batch_prepare(kmc_batch* b, double kT, int32_t precision) {
    typedef kmc_batch::Group Group;
    if (!b) return set_err(KMC_E_ARG, "null batch");
    if (precision != KMC_PREC_FP64 && precision != KMC_PREC_MIXED) return set_err(KMC_E_ARG, "bad precision");
    b->kT = kT;
    b->mixed = precision == KMC_PREC_MIXED;
    std::vector<MatArgs> mat;
    std::vector<SetupArgs> setup;
    std::vector<EvalArgs> nearA;
    std::vector<CombineArgs> comb;
    std::vector<SelectArgs> sel;
    for (Group* g : b->groups) delete g;
    b->groups.clear();
    std::vector<std::vector<EvalArgs>> far;
    std::vector<std::vector<LR2Args>> lr;
    std::vector<std::vector<SweepArgs>> sweep;
    std::vector<std::vector<TailArgs>> tail;
    std::vector<std::vector<int2>> sweep_tab;
    BatchShape all{0, 0, 0, 0, 0};
    // near field: the per-point kernel when every system selects the same lane count
    // (so a batch step sums in the order of the systems' own steps), else one lane per leaf
    // (ff-only systems have no near field and do not take part)
    bool first = true;
    b->near_lanes = 27;
    for (const kmc_batch::Sys& y : b->sys) {
        if (y.c->bc == KMC_BC_FF_ONLY) continue;
        if (first) b->near_lanes = y.c->near_lanes;
        else if (y.c->near_lanes != b->near_lanes) b->near_lanes = 27;
        first = false;
    }
    if (b->near_lanes == 0) b->near_lanes = 27;
    if (b->near_lanes < 0) b->near_lanes = -b->near_lanes;   // batches use the per-point kernels
    for (size_t i = 0; i < b->sys.size(); i++) {
        kmc_batch::Sys& y = b->sys[i];
        if (!y.set) return set_err(KMC_E_STATE, "batch system " + std::to_string(i) + " has no move stream");
        kmc_ctx* c = y.c;
        const int64_t P = y.P, T = y.P * y.k;
        int rc = ensure_eval(c, T, P);
        if (rc) return rc;
        if (uses_lr(c)) {
            rc = ensure_lr_work(c, P);
            if (rc) return rc;
        }
        if (!step_sized(c, P, T)) return set_err(KMC_E_STATE, "batch: work buffers not sized");
        const int ne = (int)(T + P);
        const bool ff = c->bc == KMC_BC_FF_ONLY;
        mat.push_back(mat_args(c, y.nsteps, P, y.k, y.pids_all, y.disp_all, y.counter, y.pid, y.es, y.new_pos));
        setup.push_back(setup_args(c, P, T, y.pid, y.es, y.new_pos, nullptr));
        // (no points for an ff-only system: its near-field blocks exit at once)
        nearA.push_back(eval_args(c, ff ? 0 : ne, c->epos4.p, c->ecell.p, c->eexcl.p, c->UD.p));
        comb.push_back(combine_args(c, P, T, y.dE, nullptr, kT, c->rates.p));
        sel.push_back(select_args(c, P, T, y.pid, y.dE, y.nsteps, y.uniform, kT, y.counter, y.selected));
        all.n++;
        all.max_ne = std::max(all.max_ne, ne);
        all.max_P = std::max(all.max_P, (int)P);
        all.max_T = std::max(all.max_T, (int)T);
        all.max_cells = std::max(all.max_cells, (int)c->ncells);
        // group by (order, periodicity, ff-only)
        const bool per = c->bc != KMC_BC_FREE_SPACE;
        size_t gi = 0;
        while (gi < b->groups.size() &&
               !(b->groups[gi]->L == c->L && b->groups[gi]->per == per && b->groups[gi]->ff == ff))
            gi++;
        if (gi == b->groups.size()) {
            Group* g = new Group();
            g->L = c->L;
            g->per = per;
            g->ff = ff;
            g->fns = c->fns;
            g->blocked = true;
            b->groups.push_back(g);
            far.emplace_back();
            lr.emplace_back();
            sweep.emplace_back();
            tail.emplace_back();
            sweep_tab.emplace_back();
        }
        Group* g = b->groups[gi];
        far[gi].push_back(eval_args(c, ne, c->epos4.p, c->ecell.p, c->eexcl.p, c->UI.p));
        // the group sweeps blocked when every system of it asks for it (and has a leaf grid
        // with a multiple of 4 per side: every batched grid has, being >= 4^3 and 2^k)
        g->blocked = g->blocked && c->sweep_block == 1 && c->ncps % 4 == 0;
        if (!ff) {
            const int nbk = (c->ncps / 4) * (c->ncps / 4) * (c->ncps / 4);
            const int si = (int)sweep[gi].size();
            for (int q = 0; q < nbk; q++) sweep_tab[gi].push_back(make_int2(si, q));
        }
        sweep[gi].push_back(sweep_rec_args(c));
        tail[gi].push_back(tail_args(c, y.counter));
        g->shape.n++;
        g->shape.max_ne = std::max(g->shape.max_ne, ne);
        g->shape.max_P = std::max(g->shape.max_P, (int)P);
        g->shape.max_T = std::max(g->shape.max_T, (int)T);
        g->shape.max_cells = std::max(g->shape.max_cells, (int)c->ncells);
        if (uses_lr(c)) {
            lr[gi].push_back(lr_args(c, P, T));
            g->lr_shape.n++;
            g->lr_shape.max_P = std::max(g->lr_shape.max_P, (int)P);
            g->lr_shape.max_T = std::max(g->lr_shape.max_T, (int)T);
        }
    }
    HIPCHK(upload(b->mat, mat));
    HIPCHK(upload(b->setup, setup));
    HIPCHK(upload(b->nearA, nearA));
    HIPCHK(upload(b->comb, comb));
    HIPCHK(upload(b->sel, sel));
    for (size_t gi = 0; gi < b->groups.size(); gi++) {
        Group* g = b->groups[gi];
        HIPCHK(upload(g->far, far[gi]));
        HIPCHK(upload(g->sweep_tab, sweep_tab[gi]));
        g->sweep_nflat = (int)sweep_tab[gi].size();
        HIPCHK(upload(g->lr, lr[gi]));
        HIPCHK(upload(g->sweep, sweep[gi]));
        HIPCHK(upload(g->tail, tail[gi]));
    }
    b->all = all;
    for (kmc_batch::Sys& y : b->sys) y.generation = y.c->generation;
    b->ready = true;
    return KMC_OK;
}

template <bool MIXED>
static void launch_batch_near(int G, int max_ne, int n, const EvalArgs* args, hipStream_t s) {
    const long long threads = (long long)max_ne * (G == 27 ? 32 : G);
    const dim3 grid((unsigned)((threads + 255) / 256), 1, n);
    switch (G) {
        case 1: hipLaunchKernelGGL((kb_near<1, MIXED>), grid, dim3(256), 0, s, args); break;
        case 4: hipLaunchKernelGGL((kb_near<4, MIXED>), grid, dim3(256), 0, s, args); break;
        case 8: hipLaunchKernelGGL((kb_near<8, MIXED>), grid, dim3(256), 0, s, args); break;
        case 16: hipLaunchKernelGGL((kb_near<16, MIXED>), grid, dim3(256), 0, s, args); break;
        case 32: hipLaunchKernelGGL((kb_near<32, MIXED>), grid, dim3(256), 0, s, args); break;
        case 64: hipLaunchKernelGGL((kb_near<64, MIXED>), grid, dim3(256), 0, s, args); break;
        default: hipLaunchKernelGGL(kb_near_cells<MIXED>, grid, dim3(256), 0, s, args); break;
    }
}

// launches fn for every group: group i on side stream (i - 1) mod nside when side
// streams are set, the first group and the rest on s; the side streams wait for s
// (event fork) and s waits for them (event join) before the next launch on s
template <class F>
static int batch_groups(kmc_batch* b, hipStream_t s, int phase, F fn) {
    const int ns = (int)b->side.size();
    if (ns == 0 || b->groups.size() < 2) {
        for (kmc_batch::Group* g : b->groups) fn(g, s);
        return KMC_OK;
    }
    HIPCHK(hipEventRecord(b->ev_fork[phase], s));
    const int used = std::min(ns, (int)b->groups.size() - 1);
    for (int i = 0; i < used; i++) HIPCHK(hipStreamWaitEvent(b->side[i], b->ev_fork[phase], 0));
    for (size_t gi = 0; gi < b->groups.size(); gi++)
        fn(b->groups[gi], gi == 0 ? s : b->side[(gi - 1) % used]);
    for (int i = 0; i < used; i++) {
        HIPCHK(hipEventRecord(b->ev_join[2 * i + phase], b->side[i]));
        HIPCHK(hipStreamWaitEvent(s, b->ev_join[2 * i + phase], 0));
    }
    return KMC_OK;
}

// average ms of one batched accept sweep of every group (the sweep launches of a batch
// step, one after another on the stream), each re-applying its systems' last accepted
// move (it changes the leaf expansions, as kmc_time_kernel's sweep does)
int kmc_batch_time_sweep(kmc_batch* b, void* stream, int32_t reps, double* h_ms) {
    if (!b || reps < 1 || !h_ms) return set_err(KMC_E_ARG, "bad arguments");
    if (!b->ready) return set_err(KMC_E_STATE, "call kmc_batch_prepare before kmc_batch_time_sweep");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++)
        for (kmc_batch::Group* g : b->groups)
            if (!g->ff)
                g->fns.bsweep(g->sweep.p, g->blocked ? g->sweep_tab.p : nullptr, g->sweep_nflat, g->shape, g->per, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    *h_ms = ms / reps;
    return KMC_OK;
}

int kmc_batch_step(kmc_batch* b, void* stream) {
    if (!b) return set_err(KMC_E_ARG, "null batch");
    if (!b->ready) return set_err(KMC_E_STATE, "call kmc_batch_prepare before kmc_batch_step");
    for (kmc_batch::Sys& y : b->sys) {
        if (y.generation != y.c->generation)
            return set_err(KMC_E_STATE, "a system changed size or buffers since kmc_batch_prepare; prepare again");
        y.c->mirrors_stale = true;
    }
    hipStream_t s = (hipStream_t)stream;
    const BatchShape& a = b->all;
    const int n = a.n;
    const int mat_threads = std::max(a.max_T, a.max_P + 1);
    hipLaunchKernelGGL(kb_materialize, dim3((mat_threads + 255) / 256, 1, n), dim3(256), 0, s, b->mat.p);
    hipLaunchKernelGGL(kb_setup, dim3((std::max(a.max_T, a.max_P) + 255) / 256, 1, n), dim3(256), 0, s, b->setup.p);
    if (b->mixed) launch_batch_near<true>(b->near_lanes, a.max_ne, n, b->nearA.p, s);
    else launch_batch_near<false>(b->near_lanes, a.max_ne, n, b->nearA.p, s);
    int rc = batch_groups(b, s, 0, [](kmc_batch::Group* g, hipStream_t t) {
        if (!g->ff) g->fns.bfar(g->far.p, g->shape, t);
        if (g->lr_shape.n > 0) g->fns.blr(g->lr.p, g->lr_shape, t);
    });
    if (rc) return rc;
    if (b->mixed)
        hipLaunchKernelGGL(kb_combine<true>, dim3((a.max_T + 63) / 64, 1, n), dim3(64), 0, s, b->comb.p);
    else
        hipLaunchKernelGGL(kb_combine<false>, dim3((a.max_T + 63) / 64, 1, n), dim3(64), 0, s, b->comb.p);
    hipLaunchKernelGGL(kb_select, dim3(1, 1, n), dim3(1024), 0, s, b->sel.p);
    rc = batch_groups(b, s, 1, [](kmc_batch::Group* g, hipStream_t t) {
        if (!g->ff)
            g->fns.bsweep(g->sweep.p, g->blocked ? g->sweep_tab.p : nullptr, g->sweep_nflat, g->shape, g->per, t);
        g->fns.btail(g->tail.p, g->shape, t);
    });
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return KMC_OK;
}
