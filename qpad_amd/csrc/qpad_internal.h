// qpad_internal.h -- device-side data layout and shared helpers of
// libqpad_hip.so (gfx950).  Layout of every slice field array in HBM:
//   f[(c * nm + k) * ld + i]   c = component, k = mode array (0 = Re m0,
//   2m-1 = Re m, 2m = Im m), i = radial node 0..nr+1, ld = padded row length
// Histories are [slice 0..nxi+1][c][k][ld] so one slice is one contiguous
// block.  Particles are structure-of-arrays, kept sorted by radial cell.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/qpad_hip.h"

namespace qpad {

constexpr int kMaxMode = 4;          // modes m = 0..4 (the most any benchmark deck uses)
constexpr int kMaxNm = 2 * kMaxMode + 1;
constexpr int kMaxSpecies = 4;
constexpr int kMaxBeams = 4;
constexpr int kSolveThreads = 512;   // workgroup of the tridiagonal solves
constexpr int kDepWaves = 4;         // default radial cells (waves) per deposit workgroup
// particles per deposit wave: a cell holding more is deposited in chunks of
// kDepChunk, the first by the cell's own wave, the others by overflow waves
// with partials of their own (rows of Species::O0/O1), so that a crowded
// cell no longer holds up a whole deposit launch
constexpr int kDepChunk = 1024;

// field kinds (source/param.f03:75-84)
enum { FK_PSI = 0, FK_EZ = 1, FK_BZ = 2, FK_BT = 3, FK_BPLUS = 4, FK_BMINUS = 5, FK_N = 6 };

__host__ __device__ inline int kre(int m) { return m == 0 ? 0 : 2 * m - 1; }
__host__ __device__ inline int kim(int m) { return 2 * m; }

struct Grid {
    int nr, nxi, M, nm, ld;
    double dr, dxi, idr, relax_fac, dt;
    int bnd, iter_max;
    double reltol, abstol;
};

struct Species {
    qpad_species_cfg cfg;
    double qbm;
    int64_t npmax;
    int64_t npp;             // host copy, refreshed after each sort
    // SoA particle state (current / alternate for the permutation)
    double *x1, *x2, *p1, *p2, *p3, *q, *gamma, *psi;
    int32_t *id;
    double *a_x1, *a_x2, *a_p1, *a_p2, *a_p3, *a_q, *a_gamma, *a_psi;
    int32_t *a_id;
    uint32_t *key, *key_alt;  // radial cell of each particle (nr = left the box)
    int32_t *cell_start;     // nr + 2 entries, cell_start[nr] = live count
    double *qn;              // neutralising background, 1 x nm x ld
    double *P0, *P1;         // per-cell deposit partials [nr][nval] (first chunk)
    // overflow chunks of crowded cells (written by the sort): rows
    // ov_start[c] .. ov_start[c+1]-1 belong to cell c, in particle order;
    // ov_desc[row] = (cell, first particle, end, 0); O0/O1 [row][nval]
    int32_t *ov_start;       // nr + 1 entries, ov_start[nr] = rows in use
    int4 *ov_desc;
    double *O0, *O1;
    int n_ov_max;            // rows allocated: npmax / kDepChunk + 1
    int *hist, *tot;         // counting-sort workspace: [chunks][nr+1], [nr+1]
    // species2d%q and its history (cbq), kept when the context tracks the
    // per-species charge (qpad_species_charge_history): 1 x nm x ld and
    // [slice 0..nxi+1][nm][ld]
    double *qsl, *h_q;
    // partial rows of the susceptibility deposit (laser contexts only):
    // [nr][nm] and [n_ov_max][nm]
    double *X0, *X1, *XO0, *XO1;
    // a neutral gas (qpad_neutral_new): the particles are the electrons
    // ionisation released, qn is the ion charge rho_ion they left
    bool neutral;
    int multi_max;           // ionisation levels followed
    double wp;               // plasma frequency (SI) of the deck's n0
    double *rate;            // [multi_max][3] ADK (C1, C2, C3)
    double *levels;          // [num_theta * nr][multi_max + 2]: level fractions,
                             // neutral residue, discrete total ion charge
    int32_t *nadd, *noff;    // [num_theta * nr + 1]: electrons released by
                             // cell in the last update, exclusive prefix
    double *add_x1, *add_x2, *add_q;   // their positions and ion charges
    int32_t *nstate;         // live before the last release, released then,
                             // released since the renew
    int32_t *nstate_host;    // pinned mirror
    void *scan_tmp;
    size_t scan_bytes;
    double *h_rho;           // rho_ion history [slice 0..nxi+1][nm][ld]
    // a non-uniform plasma (qpad_species_set_profile): the transverse
    // density of every lattice point [num_theta][nr][ppc1][ppc2] and the
    // longitudinal factor of the next injection; the kept points' exclusive
    // prefix [lattice + 1] and its scan scratch
    double *den_perp;
    double den_lon;
    int32_t *inj_keep, *inj_off;
    void *inj_tmp;
    size_t inj_tmp_bytes;
    int32_t *inj_host;       // pinned: particles the last injection kept
};

struct Beam {
    qpad_beam_cfg cfg;
    double qbm;
    int64_t npmax, npp;
    double *x1, *x2, *x3, *p1, *p2, *p3, *q;
    double *a_x1, *a_x2, *a_x3, *a_p1, *a_p2, *a_p3, *a_q;
    uint32_t *key, *key_alt, *idx, *idx_alt;
    int32_t *cell_start;     // ncell + 2 entries (cell = (mm-1) * nr + (nn-1))
    void *sort_tmp;
    size_t sort_tmp_bytes;
};

struct EventPool {
    std::vector<hipEvent_t> ev;   // pairs (start, stop)
    size_t used = 0;
    double acc_ms = 0.0;
    int64_t launches = 0;
};

struct IterState {        // device-resident iteration bookkeeping
    double rel, abs_;
    int done;             // converged in the current slice
    int counter;          // last-block-done counter of the B-perp solve
    int iters;            // iterations executed in the current slice
    int total_iters;      // iterations summed over the slices of a 3D step
    int at_max;           // slices of the 3D step that ended at iter_max
    int sub_iters;        // iterations of the current slice's earlier sub-cycles
};

}  // namespace qpad

namespace qpad {
struct SeedNet;   // the learned B_perp seed (k_seednet.hip)
}

struct qpad_ctx {
    qpad::Grid g;
    hipStream_t stream;
    int device;
    // slice fields
    double *f[QPAD_NFIELD];
    int ncomp[QPAD_NFIELD];
    // histories (slices 0..nxi+1)
    double *h_psi, *h_e, *h_b, *h_b_spe, *h_e_spe, *h_cu, *h_q_spe, *h_q_beam, *h_b_beam;
    // solver factors [kind][mode][lo|inv|cp][K * kSolveThreads], transposed
    double *fac;
    // right-hand sides of a slice's solves [4 (M+1) + 3 nm][nr]
    double *src;
    // convergence_tester vectors (nr + 2) and per-workgroup maxima
    double *conv_re, *conv_im, *conv_blk;
    qpad::IterState *it;      // device
    qpad::IterState *it_host; // pinned host mirror
    // adaptive speculation: the deposit that decides convergence writes
    // (done, iterations) into it_report (coherent pinned host memory); the
    // host waits on dec_ev recorded after it
    qpad::IterState *it_report;
    hipEvent_t dec_ev;
    int adaptive, spec_margin, spec_guess;
    int dep_cells;            // radial cells per deposit workgroup (2, 4, 8)
    int dep_mfma;             // current deposit in the matrix form (1) or the reduce-scatter form (0)
    int nspecies, nbeams;
    qpad::Species sp[qpad::kMaxSpecies];
    qpad::Beam bm[qpad::kMaxBeams];
    bool timing;
    qpad::EventPool pools[3];
    bool step_open;
    bool species_q;           // per-species charge history on (Species::qsl, h_q)
    qpad::SeedNet *net;       // null unless qpad_seednet_create
    int net_pre;              // slice whose seed is already queued (0: none)
    // laser (PGC pushers): the envelope history [(nxi + 3)][nm][ld], row
    // s + 1 = slice s (-1 .. nxi + 1), and the susceptibility history
    // [(nxi + 2)][nm][ld]; allocated by the first qpad_set_laser or PGC
    // species (laser true), after which every slice takes the envelope and
    // deposits chi
    double *lz_ar, *lz_ai, *h_chi;
    bool laser;
    // sub-cycling (simulation.algorithm "subcycling"): expansion_fac_max,
    // expansion_fac_clamped, dt_2d_min; dt2d and clamp2d are the 2D step and
    // the expansion-factor clamp of the current deposits and pushes (dxi and
    // none outside a sub-cycled slice); the last slice's sub-cycles, their
    // step and the expansion factor they came from
    bool subcyc;
    double sc_fac_max, sc_fac_clamped, sc_dt_min;
    double dt2d, clamp2d;
    double *sc_ef;            // device: the largest gamma / (gamma - p_z), as bits
    double *sc_ef_host;       // pinned mirror
    int last_nsub;
    double last_dsub, last_ef;
};

namespace qpad {

// ---- error plumbing (host) ----
void set_error(const std::string &msg);
#define QPAD_HIP_CHECK(expr)                                                 \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            ::qpad::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return QPAD_E_HIP;                                               \
        }                                                                    \
    } while (0)

#define QPAD_TRY(expr)            \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != 0) return rc_; \
    } while (0)

// p plus value v (rows of nv values) of overflow rows r0 .. r1-1, added one
// row after another (four loads in flight)
__device__ __forceinline__ double add_rows(double p, const double *O, int r0, int r1, int nv, int v)
{
    int o = r0;
    for (; o + 4 <= r1; o += 4) {
        const double x0 = O[(size_t)o * nv + v], x1 = O[(size_t)(o + 1) * nv + v];
        const double x2 = O[(size_t)(o + 2) * nv + v], x3 = O[(size_t)(o + 3) * nv + v];
        p += x0;
        p += x1;
        p += x2;
        p += x3;
    }
    for (; o < r1; o++) p += O[(size_t)o * nv + v];
    return p;
}

// deposit partial v (rows of nv values) of cell c: the row of the cell's
// own wave plus the cell's overflow rows os[c] .. os[c+1]-1, added in order
__device__ __forceinline__ double cell_partial(const double *P, const double *O, const int32_t *os, int c, int nv,
                                               int v)
{
    return add_rows(P[(size_t)c * nv + v], O, os[c], os[c + 1], nv, v);
}

// rows per thread of the field solves: the smallest power of two K with
// K * kSolveThreads >= nr (K <= 16)
inline int rows_per_thread(int nr)
{
    int k = 1;
    while (k * kSolveThreads < nr) k <<= 1;
    return k;
}

}  // namespace qpad
