// kernels.h -- host-side launch wrappers of the HIP kernels (one stream per
// context; every wrapper only enqueues work).
#pragma once

#include "qpad_internal.h"

namespace qpad {

// ---- fields (k_fields.hip) ----
// psi from the q deposit partials (+ background) and Bz of b_spe from cu;
// also 'record' of the convergence tester and reset of the iteration state.
int launch_psi_bz(qpad_ctx *c, bool do_psi, bool do_bz, bool do_record);
// psi from the already assembled q_spe (field_psi%solve)
int launch_psi_from_qspe(qpad_ctx *c);
// b = b_spe + b_beam(j), Ez of e from cu, E_perp of e from b and psi.
// final = slice epilogue: q_spe += cu_z, e_spe, history copies of slice j.
int launch_ez_et(qpad_ctx *c, int j, bool final, bool check_done);
// fine-grained field pieces of the drop-in API
int launch_ez_only(qpad_ctx *c, int dst, int jay);
int launch_et_only(qpad_ctx *c, int dst, int bsrc);
int launch_bz_only(qpad_ctx *c, int dst, int jay);
int launch_add_b(qpad_ctx *c, int j);
// cu, acu, amu from the amjdeposit partials of every species
int launch_combine_j(qpad_ctx *c, bool check_done);
// add one species' cu, acu, amu into the totals (species2d%amjdp)
int launch_combine_j_species(qpad_ctx *c, int sp);
// dcu, B_perp (B+/B-), Bz of b_spe, convergence 'compare' (last block)
int launch_bperp_bz(qpad_ctx *c, bool do_bperp, bool do_bz, bool do_compare, bool check_done);
int launch_djdxi_only(qpad_ctx *c);
// beam B_perp of every slice 1..nxi from the q_beam history
int launch_bt_beam_all(qpad_ctx *c);
// beam B_perp of slice j into the b_beam slice field
int launch_bt_beam_one(qpad_ctx *c, int j);
// convergence_tester 'compare' alone
int launch_compare_only(qpad_ctx *c);
// field kernels of predictor-corrector iteration l after its current
// deposit: cu/acu/amu and every right-hand side (k_src), the solves
// (k_solve), then b_spe, b, E_perp and the convergence maxima (k_bfields);
// each returns at once once the iteration has converged
int launch_iteration_fields(qpad_ctx *c, int j, int l);
// slice epilogue: cbq, e_spe, history copies, beam fields of slice j, the
// J_perp predictor for the next slice
int launch_epilogue(qpad_ctx *c, int j);
int slice_kernels_init();
int sort_kernels_init();   // LDS limit of the sort scatter (k_sort.hip)
// start of a slice after the charge deposit: q_spe and the right-hand
// sides, the psi, Bz and Ez solves (none reads B_perp), then the
// iteration's seed into b_spe (seed, or null to keep b_spe), 'record', b
// and E_perp for the first current deposit
int launch_slice_sources(qpad_ctx *c);
int launch_start_solves(qpad_ctx *c);
int launch_start_eperp(qpad_ctx *c, int j, const double *seed);

// learned B_perp seed (k_seednet.hip), on its own stream beside the slice:
// seednet_before_sources makes the main stream wait for slice j's seed (and
// queues its inputs and inference first if they were not queued at the end
// of slice j - 1) and returns the seed to copy into b_spe (null without one);
// seednet_after_epilogue queues slice j's Adam step (train mode) and slice
// j + 1's inputs and inference (next > 0), behind the epilogue: behind the
// point seednet_mark recorded on the main stream (marked), or behind the
// main stream's current end;
// seednet_quiesce makes the main stream wait for all of the network's work
int seednet_active(const qpad_ctx *c);
int seednet_before_sources(qpad_ctx *c, int j, const double **seed);
int seednet_mark(qpad_ctx *c);
int seednet_after_epilogue(qpad_ctx *c, int j, int next, bool marked = false);
int seednet_quiesce(qpad_ctx *c);
int seednet_step_begin(qpad_ctx *c);
void seednet_destroy(qpad_ctx *c);
// rows of the solve right-hand-side buffer
int src_rows_host(int M, int nm);
// rows of the solve right-hand-side buffer

// ---- plasma (k_plasma.hip) ----
int launch_inject(qpad_ctx *c, int sp);
// scan scratch of a profiled injection of n lattice points
int inject_scan_bytes(int64_t n, size_t *bytes);
// cell_start[c] = lower bound of c in the sorted keys, c = 0 .. ncs-1
int launch_cell_start(qpad_ctx *c, const uint32_t *key, int64_t n, int ncs, int32_t *cs);
int launch_sort(qpad_ctx *c, int sp);            // sort by key, permute, cell starts
int launch_sort_all(qpad_ctx *c);                // the same for every species in one set of launches
int launch_keys(qpad_ctx *c, int sp);            // keys of the current positions
int launch_qdeposit(qpad_ctx *c, int sp);        // per-cell partials of rho - Jz
int launch_qdeposit_range(qpad_ctx *c, int s0, int s1);   // species s0 .. s1-1 in one launch
int launch_qdeposit_all(qpad_ctx *c);            // every species in one launch
int launch_combine_q(qpad_ctx *c, int sp, double *out, double scale, bool accumulate);
// deposit partial rows of one species: P0/P1 nr rows, O0/O1 n_ov_max rows
struct DepRows {
    double *P0, *P1, *O0, *O1;
};
int launch_qdeposit_rows(qpad_ctx *c, int sp, const DepRows &rows);
int launch_combine_q_rows(qpad_ctx *c, int sp, const DepRows &rows, double *out, double scale, bool accumulate);
// the susceptibility deposit (deposit_chi) of species sp into rows
int launch_chi_deposit_rows(qpad_ctx *c, int sp, const DepRows &rows);
// l: predictor-corrector iteration of the slice (1-based; l >= 2 is a
// speculative launch that first evaluates the convergence of iteration l-1
// and returns at once when it has converged; l = 0 for a plain deposit);
// timed: bracket the launch with the timing events of pool 0
int launch_amjdeposit(qpad_ctx *c, int sp, int l, bool store_diag, bool timed, bool report = false);
// the same for every species in one launch (the workgroup rows are the
// species); report: from species 0
int launch_amjdeposit_all(qpad_ctx *c, int l, bool store_diag, bool timed, bool report = false);
// threads per workgroup of the B_perp field kernel, and its workgroups (one
// convergence maximum pair each; at most kConvMaxBlocks)
constexpr int kBfThreads = 64;
inline int conv_blocks(int nr) { return (nr + 2 + kBfThreads - 1) / kBfThreads; }
int launch_push(qpad_ctx *c, int sp);            // push_u + push_x + keys
int launch_interp_psi(qpad_ctx *c, int sp);      // psi at the particles (standard pusher only)
int launch_push_all(qpad_ctx *c);                // the same for every species in one launch
int launch_add_qn_into_qspe(qpad_ctx *c, int sp);
// per-species charge (species2d%q, cbq): q of species sp from its charge
// deposit partials (at the slice start, before the current deposit reuses
// the rows), and at the slice end its J_z added and the result stored in
// history slice j
int launch_species_q_start(qpad_ctx *c, int sp);
int launch_species_cbq(qpad_ctx *c, int sp, int j);

// ---- laser (k_laser.hip) ----
// the envelope and its gradients of slice j from the envelope history
int launch_laser_slice(qpad_ctx *c, int j);
// chi of every species (with the running rescaling) into the chi field and
// history slice j
int launch_deposit_chi(qpad_ctx *c, int j);

// ---- neutral gas (k_neutral.hip) ----
int neutral_scan_bytes(int ncells, size_t *bytes);
// every cell back to the neutral gas, no ion charge
int launch_neutral_reset(qpad_ctx *c, int sp);
// rho_ion (qn) += the ion charge of the electrons released at the end of
// the previous slice
int launch_ion_deposit(qpad_ctx *c, int sp);
// levels from the slice's final E, the released electrons appended (the
// host waits for their count, which sizes the push and the sort)
int launch_neutral_update(qpad_ctx *c, int sp);
// the ADK row (C1, C2, C3) of level `level` of element Z, or null
const double *adk_row(int element, int level);

// ---- sub-cycling (k_subcyc.hip) ----
// the largest gamma / (gamma - p_z) of any particle, at least 1 (waits for
// the stream)
int launch_exp_fac_max(qpad_ctx *c, double *out);
// fold the iterations of a sub-cycle into the step's counts (reset: at the
// start of a slice)
int launch_iter_fold(qpad_ctx *c, bool reset);
// get_subcyc_step: the sub-cycles of a slice and their 2D step
int subcyc_step(double exp_fac, double exp_fac_max, double dt, double dt_min, double *dt_sub);

// ---- beam (k_beam.hip) ----
int launch_beam_inject(qpad_ctx *c, int ib);
// lattice beam (profile "standard"): lattice_count flags and numbers the
// kept lattice points (temporary device buffers, released by lattice_free);
// lattice_write stores them into the beam, mirrors them for a quiet start,
// tilts and offsets them, and sets npp
struct LatticePlan {
    int64_t n, kept;         // lattice points in the range box, points kept
    uint32_t *flag, *pos;
    void *tmp;
    size_t tmp_bytes;
};
// a tabulated lattice density: the box xi cells k0 .. k0 + nk - 1 and r
// cells 1 .. ni, den in the lattice order (device memory)
struct LatticeTable {
    const double *den;
    int k0, nk, ni;
};
int lattice_range_box(qpad_ctx *c, const qpad_beam_std_cfg &cf, int *k0, int *nk, int *ni);
int lattice_count(qpad_ctx *c, const qpad_beam_std_cfg &cf, LatticePlan *pl, const LatticeTable *tab = nullptr);
int lattice_write(qpad_ctx *c, const qpad_beam_std_cfg &cf, const LatticePlan &pl, Beam &bm,
                  const LatticeTable *tab = nullptr);
void lattice_free(LatticePlan *pl);
int launch_beam_sort(qpad_ctx *c, int ib);
int launch_beam_keys(qpad_ctx *c, int ib);
int launch_beam_qdeposit(qpad_ctx *c);           // all beams into h_q_beam
int launch_beam_push(qpad_ctx *c, int ib);

// ---- generic ----
// device address of the static scratch of qpad_species_qdp_noalloc on the
// current device dev (noalloc.hip; once per device, from qpad_create)
int noalloc_init_device(int dev);
int field_kernels_init(size_t dyn_lds);
int sort_scratch(size_t n, int bits, size_t *bytes);
// histogram entries of the plasma counting sort for up to npmax particles
int sort_workspace(int64_t npmax, int nr, size_t *hist_ints);
int launch_fill(qpad_ctx *c, double *p, size_t n, double v);
int launch_copy(qpad_ctx *c, double *dst, const double *src, size_t n);

// event timing of dominant kernels
void time_begin(qpad_ctx *c, int which, hipEvent_t *ev_stop_out);
void time_end(qpad_ctx *c, int which, hipEvent_t ev_stop);

}  // namespace qpad
