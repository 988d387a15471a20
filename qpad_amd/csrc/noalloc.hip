// noalloc.hip -- the charge deposit with its partial rows in a static device
// array, for callers that cannot allocate device memory (code under stream
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// qpad_species_qdp deposits a species' rho - Jz into per-cell partial rows
// (P0/P1: the contributions of each radial cell to its inner and outer node;
// O0/O1: the same for the overflow chunks of crowded cells) and then sums
// them into q_spe with the axis rules and 1/r.  The rows are scratch: every
// call rewrites every row it reads.  qpad_species_qdp_noalloc runs the same
// kernels in the same order with those rows in `scratch` below, so its
// result is bit for bit the same, and calls no hipMalloc, hipFree or any
// other allocating routine.
//
// `scratch` is a host-side stand-in in host code: its device address on
// each GPU comes from hipGetSymbolAddress, taken once per device at context
// creation (qpad_create, with the context's device current) and kept in
// g_scratch[device].  Each GPU holds its own copy of the array until the
// process ends.  Every call on one GPU shares it, so calls that overlap on
// one GPU (two contexts on different streams, or host threads) race.
#include <mutex>

#include "kernels.h"
#include "qpad_internal.h"

namespace qpad {

// 2 MiB: the partial rows of the largest species the library accepts
// (nr <= 8192, modes 0..4, at most 2 * 1,048,576 particles of one species,
// so n_ov_max <= 2049): 2 * 9 * (8192 + 2049) doubles = 1.47 MiB
constexpr size_t kScratchDoubles = (size_t)2 << 17;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// device address of `scratch` on the current device (once per device)
int noalloc_init_device(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index beyond the static scratch table");
        return QPAD_E_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    QPAD_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        set_error("hipGetSymbolAddress returned a null address for the static scratch array");
        return QPAD_E_HIP;
    }
    g_scratch[dev] = (double *)p;
    return 0;
}

}  // namespace qpad

using namespace qpad;

extern "C" {

int qpad_species_qdp_noalloc(qpad_ctx *c, int32_t s)
{
    if (!c || s < 0 || s >= c->nspecies) {
        set_error("bad species");
        return QPAD_E_INVALID;
    }
    double *base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    if (!base) {
        set_error("static scratch not set up on this device (qpad_create sets it up)");
        return QPAD_E_STATE;
    }
    const Species &sp = c->sp[s];
    const size_t nm = (size_t)c->g.nm, pr = (size_t)c->g.nr * nm, orow = (size_t)sp.n_ov_max * nm;
    if (2 * pr + 2 * orow > kScratchDoubles) {
        set_error("species too large for the 2 MiB static scratch of qpad_species_qdp_noalloc "
                  "(needs 16 * nm * (nr + n_ov_max) bytes of partial rows)");
        return QPAD_E_INVALID;
    }
    const DepRows rows{base, base + pr, base + 2 * pr, base + 2 * pr + orow};
    QPAD_TRY(launch_qdeposit_rows(c, s, rows));
    QPAD_TRY(launch_combine_q_rows(c, s, rows, c->f[QPAD_F_Q_SPE], 1.0, true));
    if (sp.cfg.neutralized) QPAD_TRY(launch_add_qn_into_qspe(c, s));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
