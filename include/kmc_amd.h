/*
 * kmc_amd.h - C ABI of the MI355X (gfx950) FMM-KMC propose/accept library.
 *
 * One opaque context per system.  All functions return 0 on success and a
 * non-zero status otherwise (KMC_E_*); kmc_last_error() gives a message.
 * Pointers named d_* are device pointers (HBM), h_* are host pointers.  The
 * stream argument is a hipStream_t passed as void* (NULL = default stream).
 * A context is used from one host thread at a time.
 *
 * Each entry point cites the reference interface it replaces (paths relative to
 * the ppmd/coulomb_kmc tree).
 */
#ifndef KMC_AMD_H
#define KMC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* boundary conditions: coulomb_kmc/common.py:36-49 (BCType) */
#define KMC_BC_FREE_SPACE 0
#define KMC_BC_NEAREST    1   /* '27' */
#define KMC_BC_PBC        2
#define KMC_BC_FF_ONLY    3

/* arithmetic of the propose kernels */
#define KMC_PREC_FP64  0      /* all FP64, library sqrt/division            */
#define KMC_PREC_MIXED 1      /* FP32 reciprocal-sqrt seeds, one FP64 Newton */

/* status codes */
#define KMC_OK            0
#define KMC_E_ARG         1   /* bad argument / shape                       */
#define KMC_E_HIP         2   /* HIP runtime error                          */
#define KMC_E_STATE       3   /* call out of order (e.g. before initialise) */
#define KMC_E_BAD_MOVE    4   /* proposal outside domain or beyond max_move */
#define KMC_E_STORE       5   /* particle store inconsistent / overflow     */
#define KMC_E_UNSUPPORTED 6   /* expansion order or option not built        */

typedef struct kmc_ctx kmc_ctx;

const char* kmc_last_error(void);

/* Smallest and largest expansion order L compiled into the library. */
int32_t kmc_min_order(void);
int32_t kmc_max_order(void);

/*
 * KMCFMM.__init__  (coulomb_kmc/kmc_fmm.py:436-542).
 * n charges in a cubic box of side `extent` centred on the origin, r = number of
 * FMM levels (leaf grid 2^(r-1) per side, coulomb_kmc/kmc_fmm.py:225), L = number
 * of expansion terms, bc = KMC_BC_*, max_move as the reference (<=0: extent),
 * energy_unit multiplies every energy.  occ_headroom: extra particle slots per
 * cell above the initial maximum occupancy (the store grows when needed).
 */
int kmc_create(kmc_ctx** out, int64_t n, double extent, int32_t r, int32_t L, int32_t bc,
               double max_move, double energy_unit, int32_t occ_headroom);
int kmc_destroy(kmc_ctx* ctx);

/*
 * Positions (n x 3) and charges (n) of the charges; the index is the charge id.
 * Bins the charges into leaf cells and builds the cell occupancy / particle tables
 * (coulomb_kmc/kmc_local.py:582-824, single rank).
 */
int kmc_set_particles(kmc_ctx* ctx, const double* h_pos, const double* h_q);

/*
 * Periodic far-field operator for bc = PBC / FF_ONLY: a symmetric sparse matrix on
 * the packed (l, m>=0) (re, im) multipole vector (2*L*(L+1)/2 rows, CSR) such that
 * E_lr(M) = 0.5 * M . (G M).  Replaces ppmd's LongRangeMTL operator and dipole
 * correction (coulomb_kmc/kmc_full_long_range.py:40, 516-560, 641-692).
 */
int kmc_set_lr_operator(kmc_ctx* ctx, int64_t nnz, const double* h_data, const int64_t* h_indptr,
                        const int64_t* h_indices);

/*
 * KMCFMM.initialise (coulomb_kmc/kmc_fmm.py:947-993): builds every leaf local
 * expansion from all charges (and their 26 images for '27'/PBC) on the GPU (see
 * kmc_set_init_method), the level-0 multipole M, and the energy.  *h_energy receives
 * the energy.
 */
int kmc_initialise(kmc_ctx* ctx, void* stream, double* h_energy);

/*
 * How kmc_initialise builds the leaf expansions (the reference takes them from ppmd's
 * FMM, coulomb_kmc/kmc_fmm.py:463-476, 956-959): 1 = direct P2L of every charge into
 * every well-separated leaf (exact; O(n x leaves)); 2 = hierarchical: multipoles, M2M
 * and M2L on the levels above the leaves' grandparents, L2L down, plus direct P2L of
 * the charges in the 12^3 leaves around each leaf's grandparent block (M2L truncation
 * at order L; needs >= 8^3 leaves); 0 = automatic (hierarchical for large systems).
 */
int kmc_set_init_method(kmc_ctx* ctx, int32_t method);

/* Leaf expansions in the reference layout [ncells][2*L*L] (re block then im block,
 * index l*l+l+m; coulomb_kmc/kmc_octal.py:61-62). */
int kmc_get_expansions(kmc_ctx* ctx, double* h_out);
int kmc_set_expansions(kmc_ctx* ctx, const double* h_in);
/* M (= E) in the reference layout [2*L*L] (coulomb_kmc/kmc_full_long_range.py:42-43). */
int kmc_get_multipole(kmc_ctx* ctx, double* h_out);
/* Cell tables: occ[ncells], ids[ncells*occmax], pos4[ncells*occmax*4]; *occmax out. */
int kmc_get_store(kmc_ctx* ctx, int32_t* h_occ, int32_t* h_ids, double* h_pos4, int32_t* h_occmax);
int kmc_store_occmax(kmc_ctx* ctx, int32_t* out);
int kmc_get_particles(kmc_ctx* ctx, double* h_pos, int32_t* h_cells);
int kmc_get_energy(kmc_ctx* ctx, double* h_energy);
int kmc_set_energy(kmc_ctx* ctx, double energy);
int kmc_ncells(kmc_ctx* ctx, int64_t* out);

/*
 * Proposal energies, tuple form (KMCFMM.propose -> _propose, coulomb_kmc/kmc_fmm.py:
 * 774-787, 1001-1105, and the diff_lib combine :583-663).  P moving charges d_pid[P];
 * their proposals are d_new_pos[d_es[p] .. d_es[p+1]) (T = es[P] rows of 3 doubles);
 * d_new_q optional new charges (NULL: unchanged).  Writes d_dE[T] = energy_unit *
 * (U1D + U1I - U0D - U0I - USI).  T must equal es[P].  Returns KMC_E_BAD_MOVE when a
 * proposal lies outside the domain or beyond max_move (kmc_mpi_decomp.py:312-345, 730).
 */
int kmc_propose(kmc_ctx* ctx, void* stream, int64_t P, int64_t T, const int64_t* d_pid,
                const int64_t* d_es, const double* d_new_pos, const double* d_new_q, double* d_dE,
                int32_t precision);

/*
 * kmc_propose without allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as kmc_propose, but the evaluation arrays live in
 * a fixed device array of KMC_NOALLOC_SCRATCH_BYTES that the library holds on each GPU
 * it runs on until the process ends; the call makes no hipMalloc / hipFree.  Returns
 * KMC_E_ARG (the message gives the bytes needed) when P moving charges and T proposals
 * need more scratch than that: (T + P)(48 + 8 f) + 4 T + 8 P bytes (f = 2 far-field
 * partials per point up to L = 16, 4 above), plus
 * 96 T + 96 P (nb + r) for pbc / ff-only (nb = 22 at L = 12, 36 at L = 16; r = 4 below
 * 4096 moving charges, else 1), plus 8 (T + P) + 132 nt + 4 when the leaf grid has a
 * multiple of 4 leaves per side (nt = leaves / 64 tiles, for the tile-sorted near
 * field), each array rounded up to 256 bytes: 3.9 MiB for 1024 charges x 8 proposals
 * at L = 12 on 32^3 leaves (pbc), 19.6 MiB for 4096 x 8 at L = 16 on 16^3 leaves.  Every call on one GPU shares that GPU's array (the
 * last call's contents stay in it): calls that overlap on one GPU (several host
 * threads or streams) race.  Synchronises the stream (bad-move check).
 */
#define KMC_NOALLOC_SCRATCH_BYTES ((int64_t)32 << 20)
int kmc_propose_noalloc(kmc_ctx* ctx, void* stream, int64_t P, int64_t T, const int64_t* d_pid,
                        const int64_t* d_es, const double* d_new_pos, const double* d_new_q, double* d_dE,
                        int32_t precision);

/*
 * KMCFMM.propose_with_dats (coulomb_kmc/kmc_fmm.py:665-762 with setup_move,
 * coulomb_kmc/kmc_mpi_decomp.py:501-616).  Device arrays: site_max_counts[nsites],
 * current_sites[n], prop_positions[n][max_prop][3], prop_masks[n][max_prop],
 * prop_charges[n][max_prop] or NULL; writes prop_energy_diffs[n][max_prop] at every
 * masked slot taken (mask > 0, at most site_max_counts[site] per charge).
 * *h_total_moves (optional) receives the number of proposals evaluated.  A
 * current_sites[i] outside [0, nsites) returns KMC_E_ARG (checked on the device).
 */
int kmc_propose_with_dats(kmc_ctx* ctx, void* stream, int64_t max_prop, const int64_t* d_site_max_counts,
                          int64_t nsites, const int64_t* d_current_sites, const double* d_prop_positions,
                          const int64_t* d_prop_masks, const double* d_prop_charges,
                          double* d_prop_energy_diffs, int32_t precision, int64_t* h_total_moves);

/*
 * KMCFMM.accept (coulomb_kmc/kmc_fmm.py:789-912): moves charge `pid` to new_pos.
 * Computes the new energy with a one-proposal FP64 propose, updates the cell tables
 * (extract/insert, coulomb_kmc/kmc_local.py:180-370), sweeps every leaf local
 * expansion (accept_local_exp, coulomb_kmc/kmc_octal.py:977-1227) and the level-0
 * multipole (coulomb_kmc/kmc_full_long_range.py:270-316).  *h_energy (optional)
 * receives the new energy.
 */
int kmc_accept(kmc_ctx* ctx, void* stream, int64_t pid, const double* h_new_pos, double* h_energy);

/*
 * The accept of KMCFMM._accept's move record (coulomb_kmc/kmc_fmm.py:811-836):
 * movedata[10] int64, words 0..6 the doubles old position, new position, charge, then
 * the charge's id (the local id here), old leaf and new leaf (linear, x fastest).  The
 * record must match the charge's current position, charge and leaf and the new
 * position's leaf; then as kmc_accept.
 */
int kmc_accept_movedata(kmc_ctx* ctx, void* stream, const int64_t* h_movedata, double* h_energy);

/*
 * One whole KMC step on the device, no host synchronisation (graph capturable
 * once the context's work buffers have been sized by a first call):
 * proposals T from d_pid/d_es/d_new_pos as kmc_propose, then Metropolis selection
 * with rates min(1, exp(-dE/kT)) and cumulative sum in proposal order against
 * uniform d_uniform[*d_counter % nsteps] (the example harness's rule,
 * examples/propose_with_dats_example.py:316-411), then the accept of the selected
 * proposal.  d_selected[*d_counter % nsteps] receives the selected proposal index
 * and *d_counter is incremented.  d_dE (T) receives the proposal energies.
 */
int kmc_step(kmc_ctx* ctx, void* stream, int64_t P, int64_t T, const int64_t* d_pid, const int64_t* d_es,
             const double* d_new_pos, double* d_dE, int64_t nsteps, const double* d_uniform, double kT,
             int64_t* d_counter, int64_t* d_selected, int32_t precision);

/*
 * Synthetic move stream for benchmarks and tests (no reference counterpart: the
 * reference's harness builds proposals on the host).  For step = *d_counter % nsteps:
 * pid = d_pids_all[step*P + p], new_pos[t] = wrap/clip(pos[pid] + d_disp_all[step*T + t])
 * with t in p*k .. p*k+k-1.  Fills d_pid[P], d_es[P+1] (k per charge), d_new_pos[T].
 * d_pids_all holds nsteps*P ids, d_disp_all nsteps*T*3 doubles.
 */
int kmc_materialize_moves(kmc_ctx* ctx, void* stream, int64_t nsteps, int64_t P, int64_t k,
                          const int64_t* d_pids_all, const double* d_disp_all, const int64_t* d_counter,
                          int64_t* d_pid, int64_t* d_es, double* d_new_pos);

/*
 * nrun steps of the synthetic move stream in one call (no reference counterpart):
 * each is kmc_materialize_moves (nsteps_moves table rows) followed by kmc_step
 * (nsteps_uniform uniforms), with the same kernels and the same results.  With
 * pipelined != 0 (not for ff-only) the accept sweep of step i runs on a second side
 * stream beside the set-up, near field and long range of step i + 1, which read no
 * leaf expansion; step i + 1's far field waits for it, and the call's stream waits
 * for the last sweep before returning (so a captured call is joined).  pipelined = 2:
 * the near field waits for the sweep too (only set-up and long range overlap it).
 * pipelined = 3: as 2, with step i + 1's far field on the sweep's stream right behind
 * the sweep, beside the near field.  pipelined = 4: the call's stream carries every
 * step's near field, combine, select and sweep; a side stream the accept tail and,
 * behind it, the next step's materialize, set-up and long range; a second side stream
 * the far field (after the set-up and the previous sweep).  Other values: KMC_E_ARG.
 * Needs the work buffers sized by a kmc_propose of the same T.
 */
int kmc_run_steps(kmc_ctx* ctx, void* stream, int64_t nrun, int64_t nsteps_moves, int64_t P, int64_t k,
                  const int64_t* d_pids_all, const double* d_disp_all, int64_t nsteps_uniform,
                  const double* d_uniform, double kT, int64_t* d_counter, int64_t* d_selected, int64_t* d_pid,
                  int64_t* d_es, double* d_new_pos, double* d_dE, int32_t precision, int32_t pipelined);

/*
 * KMCFMM.eval_field (coulomb_kmc/kmc_fmm.py:544-567): potential of all charges at n
 * points d_points[n][3] (device) into d_out[n]: near field over the 27 leaves around
 * the point's leaf (coulomb_kmc/kmc_local.py:1085-1182), that leaf's local expansion
 * (coulomb_kmc/kmc_octal.py:1417-1438) and, for pbc / ff-only, the long-range term
 * (coulomb_kmc/kmc_full_long_range.py:341-387).  energy_unit is not applied.
 * Synchronises the stream before returning.
 */
int kmc_eval_field(kmc_ctx* ctx, void* stream, int64_t n, const double* d_points, double* d_out);

/*
 * Charge count of the context and a new one for the next kmc_set_particles (the count
 * changes with kmc_extract / kmc_inject; set it before re-initialising with another
 * set of charges).  kmc_set_particle_count drops the particles: set them and
 * initialise again.
 */
int kmc_particle_count(kmc_ctx* ctx, int64_t* out);
int kmc_set_particle_count(kmc_ctx* ctx, int64_t n);

/*
 * Energy of each charge d_ids[n] (device, local ids) with all others, energy_unit
 * applied: near field without the charge itself, its leaf's local expansion and, for
 * pbc / ff-only, the long-range term (KMCFMM.get_energy,
 * coulomb_kmc/kmc_inject_extract.py:82-124, per charge before its row sum; parts from
 * kmc_local.py:374-396, kmc_octal.py:363-375, kmc_full_long_range.py:123-140).
 * Synchronises the stream.
 */
int kmc_charge_energies(kmc_ctx* ctx, void* stream, int64_t n, const int64_t* d_ids, double* d_out);
/*
 * Potential at each charge d_ids[n] from all the other charges (the same parts as
 * kmc_charge_energies at unit charge, energy_unit not applied): the per-charge
 * potential that KMCFMM.initialise has ppmd's FMM write into group._kmc_fmm_potential
 * (coulomb_kmc/kmc_fmm.py:942-958); 0.5 sum_i q_i phi_i energy_unit is the energy.
 * Synchronises the stream.
 */
int kmc_charge_potentials(kmc_ctx* ctx, void* stream, int64_t n, const int64_t* d_ids, double* d_out);

/*
 * Energy of each of ngroups groups of m charges alone in the domain, energy_unit
 * applied (KMCFMM.compute_energy, coulomb_kmc/kmc_inject_extract.py:145-165, with the
 * direct sums of coulomb_kmc/kmc_direct.py:150-300: pairs; '27' / pbc: half of every
 * pair and self pair over the 26 neighbour images; pbc / ff-only: 0.5 M.(G M) of the
 * group's multipole).  Device d_pos[ngroups*m][3], d_q[ngroups*m] -> d_out[ngroups].
 */
int kmc_group_energy(kmc_ctx* ctx, void* stream, int64_t ngroups, int64_t m, const double* d_pos,
                     const double* d_q, double* d_out);

/*
 * KMCFMM.propose_extract (coulomb_kmc/kmc_inject_extract.py:178-213): for each row of
 * d_ids[nrows][m] (device local ids) the energy change of removing the row's charges:
 * (self_energy ? energy of the row alone : 0) - sum of the row's charge energies.
 * d_out[nrows].  Synchronises the stream.
 */
int kmc_propose_extract(kmc_ctx* ctx, void* stream, int64_t nrows, int64_t m, const int64_t* d_ids,
                        int32_t self_energy, double* d_out);

/*
 * KMCFMM.propose_inject (coulomb_kmc/kmc_inject_extract.py:216-237): energy change of
 * adding n charges d_pos[n][3], d_q[n] (device): their energy alone plus
 * energy_unit * sum q_i phi(x_i).  d_out[0].  Synchronises the stream.
 */
int kmc_propose_inject(kmc_ctx* ctx, void* stream, int64_t n, const double* d_pos, const double* d_q,
                       double* d_out);

/*
 * KMCFMM.extract (coulomb_kmc/kmc_inject_extract.py:240-301): removes the k charges
 * h_ids[k] (host, unique local ids; not every charge): energy += propose_extract of
 * the ids as one row, the leaf expansions and the level-0 multipole lose the charges
 * (kmc_octal.py:1091-1217 extraction mode, kmc_full_long_range.py:270-284), the cell
 * tables drop them (kmc_local.py:180-271).  Ids stay 0 .. n-k-1: taking the ids in
 * descending order, the charge with the highest id moves into each freed id.
 * *h_energy (optional) receives the new energy.  Changes the charge count.
 */
int kmc_extract(kmc_ctx* ctx, void* stream, int64_t k, const int64_t* h_ids, double* h_energy);

/*
 * KMCFMM.inject (coulomb_kmc/kmc_inject_extract.py:304-424): adds k charges h_pos[k][3]
 * (inside the domain), h_q[k] (host) with ids n .. n+k-1: energy += propose_inject,
 * cell tables (kmc_local.py:274-359), leaf expansions (kmc_octal.py:1091-1217
 * injection mode) and level-0 multipole (kmc_full_long_range.py:287-303).  Grows the
 * cell tables and particle arrays when needed.  Changes the charge count.
 */
int kmc_inject(kmc_ctx* ctx, void* stream, int64_t k, const double* h_pos, const double* h_q, double* h_energy);

/*
 * DiscoverInjectExtract (coulomb_kmc/kmc_inject_extract.py:439-533), context free:
 * d_inject_flags[s] = number of charges d_pos[n][3] within squared distance tol2 of
 * inject site s; d_extract_flags[i] = 1 + the first extract site within tol2 of charge
 * i, 0 if none (all device arrays).  The reference compares squared distances with
 * 1e-5.  Synchronises the stream.
 */
int kmc_discover_sites(void* stream, int64_t n, const double* d_pos, int64_t n_inject,
                       const double* d_inject_sites, int64_t n_extract, const double* d_extract_sites,
                       double tol2, int64_t* d_inject_flags, int64_t* d_extract_flags);

/*
 * One lattice KMC step on the device, Algorithm 2 of the reference's examples
 * (docs/source/example_bookkeeping.rst; examples/propose_with_dats_example.py:217-450):
 * every charge proposes position + h_offsets[m] (host [M][3], each component within
 * max_move), wrapped into the domain (wrap != 0; else proposals outside the domain
 * get mask 0) -> d_new_pos[N*M][3], d_masks[N*M] = 1; proposals within squared
 * distance tol2 (<= leaf width^2; the example's TOL is 0.01) of another charge get
 * mask 0 (minimum image when wrapping); all N*M proposals are evaluated -> d_dE;
 * rates exp(-dE/kT) (rate_rule 1, the example's rate kernel) or min(1, exp(-dE/kT))
 * (rate_rule 0), 0 where masked; selection and accept as kmc_step, with proposal index
 * i*M + m written to d_selected[*d_counter % nsteps].  The first call for a charge count
 * and offset set sizes the work buffers (eager); later calls allocate nothing and may be
 * captured in a graph.  No host synchronisation.
 */
int kmc_lattice_step(kmc_ctx* ctx, void* stream, int64_t M, const double* h_offsets, double tol2, int32_t wrap,
                     double* d_new_pos, int64_t* d_masks, double* d_dE, int64_t nsteps, const double* d_uniform,
                     double kT, int32_t rate_rule, int64_t* d_counter, int64_t* d_selected, int32_t precision);

/*
 * Flags raised by device-side steps (kmc_step, kmc_run_steps, kmc_lattice_step, batches),
 * or-ed into *h_flags and cleared by the call: 1 the moved charge was not found exactly
 * once in its leaf's table, 2 the selected move's target leaf was full (not accepted;
 * kmc_store_reserve before the run rules it out), 4 no proposal had a positive rate, 8 a moving-charge id out of range,
 * 16 a synthetic move-stream row out of range, 32 a proposal outside the domain or
 * beyond max_move, or malformed es (nothing accepted in that step; the host propose
 * paths return KMC_E_BAD_MOVE for the same inputs).
 */
int kmc_check_flags(kmc_ctx* ctx, void* stream, int32_t* h_flags);

/*
 * Room in the cell tables for device-side steps.  Replaces the growth of the particle
 * store inside accept (LocalParticleData._resize_particle_store,
 * coulomb_kmc/kmc_local.py:558-578, called at :321-330 when the target cell is full):
 * device steps cannot allocate, so between runs the caller reserves room for the next
 * run.  Every leaf gets at least `slots` free slots (a run of n steps adds at most n
 * charges to any leaf, so slots = n makes flag 2 impossible in that run); when the
 * fullest leaf has fewer, the tables grow, contents kept, to its occupancy + slots +
 * the context's occupancy headroom, and *h_grew is 1: graphs captured before then hold
 * the old tables and must be captured again (batches: kmc_batch_prepare again).
 * Synchronises the stream (not capturable).
 */
int kmc_store_reserve(kmc_ctx* ctx, void* stream, int32_t slots, int32_t* h_grew);

/*
 * Tile choices of the hot kernels (filled by the autotuner in coulomb_kmc_amd/tune.py):
 * near_lanes = lanes per evaluation point in the near-field kernel (1,4,8,16,32,64, or
 * 27 = one lane per neighbour leaf); accept sweep: sweep_pairs = column pairs
 * (m, L-1-m) per lane (1, 2), sweep_block = leaves (threads) per workgroup (64, 128,
 * 192, 256), or 0: fused (64 consecutive leaves x all column groups per workgroup,
 * update staged through LDS), or 1: blocked (a 4x4x4 block of leaves per workgroup,
 * orders of each source image truncated where their sum changes the potential in the
 * block by less than 2^-60 per unit charge; accepts only, leaf grids with a multiple of
 * 4 leaves per side, else as 0), sweep_split = waves sharing one leaf's 27 images (1,
 * or 3: one image layer each, summed in LDS; workgroups of 64 leaves x 3; '27'/PBC only).
 * near_lanes -4 / -8 / -16: points counting-sorted by 4x4x4 leaf tile, each tile's 6^3
 * leaves of charges staged in LDS once, G lanes per point (propose paths; grids with a
 * multiple of 4 leaves per side, else G lanes per point from memory).
 */
int kmc_set_tiles(kmc_ctx* ctx, int32_t near_lanes, int32_t sweep_block, int32_t sweep_pairs,
                  int32_t sweep_split);
/*
 * Order of the near field's evaluation entries (a tile choice, with kmc_set_tiles):
 * 0 the proposals' order; 1 counting-sorted by 4x4x4 leaf tile, each XCD taking one
 * contiguous range of the sorted entries, for the per-point kernels (near_lanes 1, 4, 8,
 * 16, 32, 64; propose paths; leaf grids with a multiple of 4 per side, else order 0).
 * The sums equal order 0's bit for bit.
 */
int kmc_set_near_order(kmc_ctx* ctx, int32_t order);

/*
 * Per-kernel profile of host-called proposes and accepts (the reference's counters:
 * c_direct_new_time / c_direct_old_time / c_direct_*_div_count,
 * coulomb_kmc/kmc_local.py:466-492; c_indirect_time / c_indirect_exec_count,
 * kmc_octal.py:352-358; c_accept_time, kmc_octal.py:176-183).  kmc_profile_enable
 * (on = 1 / 0) clears the sums and turns the recording on / off: HIP event pairs
 * around the near field, far field, long range and accept sweep of every call that is
 * not under graph capture (kmc_propose_noalloc records none), and a count of the
 * near field's pair evaluations.  kmc_profile_read (synchronises) writes h_out[9]:
 * near-field ms, its evaluation points, its pair evaluations; far-field ms, points;
 * long-range ms, proposals; accept-sweep ms, sweeps.  The near field evaluates the
 * new and old positions in one launch (the reference's direct_new and direct_old).
 */
int kmc_profile_enable(kmc_ctx* ctx, int32_t on);
int kmc_profile_read(kmc_ctx* ctx, double* h_out);

/* Accept sweep only, for a hop of charge q from h_old to h_new (timing / tests). */
int kmc_sweep_hop(kmc_ctx* ctx, void* stream, double q, const double* h_old_pos, const double* h_new_pos);

/*
 * Measurement hook (no reference counterpart): average duration in ms of `reps`
 * back-to-back launches of one kernel, timed with HIP events on `stream`:
 * which = 0 accept sweep (re-applies the last accepted move: changes the state,
 * call last), 1 near field, 2 far field, 3 long-range proposal term (on the
 * evaluation arrays of the last propose).
 */
int kmc_time_kernel(kmc_ctx* ctx, void* stream, int32_t which, int32_t reps, int32_t precision, double* h_ms);

/* Measurement hook (no reference counterpart): the roofline denominators measured on
 * the device, beside the published peaks.  Allocates two buffers of
 * `bytes`, times `reps` launches of a streaming copy and of an FP64 FMA loop on
 * `stream`; h_out[0] = copy GB/s (read + write bytes / time), h_out[1] = FP64
 * TFLOP/s (FMA = 2 flops).  bytes in [1 MiB, 16 GiB]. */
int kmc_probe_peaks(void* stream, int64_t bytes, int32_t reps, double* h_out);

/*
 * Batches of independent systems (no reference counterpart: the reference steps one
 * system per process).  The step of every system in the batch is one launch per kernel
 * kind (per expansion order for the order-specialised kernels), the system being
 * blockIdx.z of a table of per-system kernel arguments.  Systems: initialised contexts
 * with >= 4^3 leaves, of any boundary condition (ff-only systems form groups of their
 * own that run the long range and the cell-table tail only, as their single steps do).  Per system, the same device move stream as
 * kmc_materialize_moves + kmc_step (pids_all[nsteps*P], disp_all[nsteps*P*k*3],
 * uniform[nsteps], counter, selected[nsteps], and pid[P], es[P+1], new_pos[P*k*3],
 * dE[P*k] work arrays).  kmc_batch_prepare sizes every context's work buffers and
 * uploads the argument tables (it allocates; call it again after host-side accepts
 * or anything else that may reallocate a context's tables); kmc_batch_step allocates
 * nothing and can be captured in a graph.  The accept sweep of each (order,
 * periodicity) group uses the blocked layout with order truncation when every system
 * of the group is set to sweep_block 1 (one flat grid over the groups' (system, 4x4x4
 * block) pairs, so no workgroup goes to a block a smaller system lacks), else the fused
 * layout with two column pairs per lane; the near field uses G lanes per point when
 * every system selects the same G (near_lanes G or -G), else one lane per leaf.
 */
typedef struct kmc_batch kmc_batch;
int kmc_batch_create(kmc_batch** out, kmc_ctx* const* ctxs, int32_t n);
int kmc_batch_destroy(kmc_batch* batch);
int kmc_batch_set_system(kmc_batch* batch, int32_t i, int64_t nsteps, int64_t P, int64_t k, const int64_t* d_pids_all,
                         const double* d_disp_all, const double* d_uniform, int64_t* d_counter, int64_t* d_selected,
                         int64_t* d_pid, int64_t* d_es, double* d_new_pos, double* d_dE);
int kmc_batch_prepare(kmc_batch* batch, double kT, int32_t precision);
int kmc_batch_step(kmc_batch* batch, void* stream);
/* Average ms of the batch's accept-sweep launches (no reference counterpart; the
 * measurement of the batched sweep): one launch per (order, periodicity) group, run
 * reps times one after another on the stream, each re-applying every system's last
 * accepted move (the leaf expansions change, as with kmc_time_kernel's sweep). */
int kmc_batch_time_sweep(kmc_batch* batch, void* stream, int32_t reps, double* h_ms);
/* Side streams of a batch (no reference counterpart): with nstreams > 1, the launches
 * that run once per (expansion order, periodicity) group -- far field, long range,
 * accept sweep and cell-table tail -- go to nstreams - 1 library-owned side streams
 * (default priority) forked from and joined back to the step's stream with events, so
 * the groups overlap and a captured step keeps its dependencies.  1 (the default):
 * every launch on the step's stream.  nstreams in [1, 16]; call outside capture. */
int kmc_batch_set_streams(kmc_batch* batch, int32_t nstreams);

#ifdef __cplusplus
}
#endif
#endif
