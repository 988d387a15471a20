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

constexpr int kMaxMode = 4;          // modes m = 0..4
constexpr int kMaxNm = 2 * kMaxMode + 1;
constexpr int kMaxSpecies = 4;
constexpr int kMaxBeams = 4;
constexpr int kSolveThreads = 256;   // workgroup of the tridiagonal solves
constexpr int kDepWaves = 4;         // waves (cells) per deposit workgroup

// field kinds (source/param.f03:72-80)
enum { FK_PSI = 0, FK_EZ = 1, FK_BZ = 2, FK_BT = 3, FK_BPLUS = 4, FK_BMINUS = 5, FK_N = 6 };

__host__ __device__ inline int kre(int m) { return m == 0 ? 0 : 2 * m - 1; }
__host__ __device__ inline int kim(int m) { return 2 * m; }

struct Grid {
    int nr, nxi, M, nm, ld;
    double dr, dxi, idr, relax_fac, dt;
    int bnd, iter_max;
    double reltol, abstol;
};

// Thomas factors of one matrix: lo = sub-diagonal, inv = 1/pivot,
// cp = super-diagonal/pivot, rows 0..nr-1 = QPAD grid points 1..nr.
struct Factor {
    const double *lo, *inv, *cp;
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
    uint32_t *key, *key_alt, *idx, *idx_alt;
    int32_t *cell_start;     // nr + 2 entries, cell_start[nr] = npp
    double *qn;              // neutralising background, 1 x nm x ld
    double *P0, *P1;         // per-cell deposit partials [nr][nval]
    void *sort_tmp;
    size_t sort_tmp_bytes;
};

struct Beam {
    qpad_beam_cfg cfg;
    double qbm;
    int64_t npmax, npp;
    double *x1, *x2, *x3, *p1, *p2, *p3, *q;
    double *a_x1, *a_x2, *a_x3, *a_p1, *a_p2, *a_p3, *a_q;
    uint32_t *key, *key_alt, *idx, *idx_alt;
    int32_t *cell_start;     // ncell + 1 entries (cell = (mm-1) * nr + (nn-1))
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
    int pad[3];
};

}  // namespace qpad

struct qpad_ctx {
    qpad::Grid g;
    hipStream_t stream;
    int device;
    // slice fields
    double *f[QPAD_NFIELD];
    int ncomp[QPAD_NFIELD];
    // histories (slices 0..nxi+1)
    double *h_psi, *h_e, *h_b, *h_b_spe, *h_e_spe, *h_cu, *h_q_spe, *h_q_beam, *h_b_beam;
    // beam q scratch for the per-beam deposit
    // solver factors [kind][mode][3][nr]
    double *fac;
    // convergence_tester vectors (nr + 2)
    double *conv_re, *conv_im;
    qpad::IterState *it;      // device
    qpad::IterState *it_host; // pinned host mirror
    int nspecies, nbeams;
    qpad::Species sp[qpad::kMaxSpecies];
    qpad::Beam bm[qpad::kMaxBeams];
    bool timing;
    qpad::EventPool pools[3];
    bool step_open;
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

inline const double *fac_ptr(const qpad_ctx *c, int kind, int m, int which)
{
    return c->fac + (((size_t)kind * (kMaxMode + 1) + m) * 3 + which) * c->g.nr;
}
inline Factor factor(const qpad_ctx *c, int kind, int m)
{
    return Factor{fac_ptr(c, kind, m, 0), fac_ptr(c, kind, m, 1), fac_ptr(c, kind, m, 2)};
}

// ---- device helpers ----

// Composition of affine maps y -> A*y + B over the threads of a workgroup.
// Returns, for thread t, the B-component of (map_{t-1} o ... o map_0), i.e.
// the value entering thread t's segment when the first segment starts from 0.
// `red` must hold 2 * (blockDim.x / 64) doubles of LDS.
__device__ inline double block_affine_exclusive(double A, double B, double *red)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwave = blockDim.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        double Ap = __shfl_up(A, d, 64);
        double Bp = __shfl_up(B, d, 64);
        if (lane >= d) {
            B = A * Bp + B;
            A = A * Ap;
        }
    }
    if (lane == 63) {
        red[2 * wave] = A;
        red[2 * wave + 1] = B;
    }
    __syncthreads();
    // prefix of the earlier waves (applied before this wave's maps)
    double PA = 1.0, PB = 0.0;
    for (int w = 0; w < wave; w++) {
        double wa = red[2 * w], wb = red[2 * w + 1];
        PB = wa * PB + wb;
        PA = wa * PA;
    }
    (void)nwave;
    // inclusive result for this thread: (A,B) o (PA,PB)
    double incB = A * PB + B;
    // exclusive: value from the previous lane, or the wave prefix for lane 0
    double exB = __shfl_up(incB, 1, 64);
    if (lane == 0) exB = PB;
    __syncthreads();
    return exB;
}

// Solve the tridiagonal system with precomputed Thomas factors in place on
// `d` (n rows, LDS or global), with the whole workgroup.  Each thread owns a
// contiguous segment; segment maps are combined with a block scan and each
// segment is then recomputed serially with Thomas' recurrences.
__device__ inline void block_thomas(const Factor f, double *d, int n, double *red)
{
    const int T = blockDim.x;
    const int t = threadIdx.x;
    const int K = (n + T - 1) / T;
    const int s0 = t * K;
    const int s1 = min(s0 + K, n);
    // forward: y_i = (d_i - lo_i y_{i-1}) inv_i  =>  y_i = al_i y_{i-1} + be_i
    double A = 1.0, B = 0.0;
    for (int i = s0; i < s1; i++) {
        double inv = f.inv[i];
        double al = -f.lo[i] * inv;
        B = al * B + d[i] * inv;
        A = al * A;
    }
    double y = block_affine_exclusive(A, B, red);
    for (int i = s0; i < s1; i++) {
        y = (d[i] - f.lo[i] * y) * f.inv[i];
        d[i] = y;
    }
    __syncthreads();
    // backward: x_i = y_i - cp_i x_{i+1}; thread order reversed
    const int tr = T - 1 - t;
    const int r0 = tr * K;
    const int r1 = min(r0 + K, n);
    A = 1.0;
    B = 0.0;
    for (int i = r1 - 1; i >= r0; i--) {
        double al = -f.cp[i];
        B = al * B + d[i];
        A = al * A;
    }
    double x = block_affine_exclusive(A, B, red);
    // x is the value of x_{r1} entering from above (0 beyond the last row)
    for (int i = r1 - 1; i >= r0; i--) {
        x = d[i] - f.cp[i] * x;
        d[i] = x;
    }
    __syncthreads();
}

}  // namespace qpad
