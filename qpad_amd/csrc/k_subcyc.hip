// k_subcyc.hip -- QPAD's sub-cycling algorithm (proj_subcyc, simulation
// algorithm "subcycling"): a slice whose plasma electrons move nearly with
// the light (gamma / (gamma - p_z) large) is pushed in several 2D steps.
// Per slice (simulation_subcyc_class.f03:230-390): the largest expansion
// factor gamma / (gamma - p_z) of any particle at the slice start
// (get_exp_fac_max, part2d_subcyc_class.f03:27-44, at least 1) sets the
// sub-cycles n and their step dxi / n (get_subcyc_step, :431-449); each
// sub-cycle repeats the charge deposit, the field solves, the
// predictor-corrector iteration with the step dxi / n and the push, whose
// push_u is followed by clamp_exp_fac (part2d_subcyc_class.f03:46-66, fused
// into k_push).  The host reads the factor back (8 bytes) to queue the
// sub-cycles.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kernels.h"

namespace qpad {

struct ExpFacArgs {
    int nr;
    const double *gamma[kMaxSpecies], *p3[kMaxSpecies];
    const int32_t *cs[kMaxSpecies];
    unsigned long long *out;   // bits of a positive double: ordered like the values
};

// workgroup rows by species (blockIdx.y), grid-stride over the live
// particles; max(1, gamma / (gamma - p_z)) per workgroup, then one atomic
// max on the bits (every value is >= 1, so the bit order is the value order)
__global__ __launch_bounds__(256) void k_exp_fac_max(const ExpFacArgs a)
{
    const int s = blockIdx.y;
    const int64_t npp = a.cs[s][a.nr];
    const double *g = a.gamma[s], *pz = a.p3[s];
    double m = 1.0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npp; p += (int64_t)gridDim.x * blockDim.x) {
        const double gp = g[p];
        const double f = gp / (gp - pz[p]);
        if (f > m) m = f;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double v = __shfl_xor(m, o);
        if (v > m) m = v;
    }
    __shared__ double wmax[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wmax[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++)
            if (wmax[w] > m) m = wmax[w];
        atomicMax(a.out, (unsigned long long)__double_as_longlong(m));
    }
}

// the iteration bookkeeping of a sub-cycle that is not the slice's last
// (reset: start of a slice)
__global__ void k_iter_fold(IterState *it, int iter_max, int reset)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (reset) {
        it->sub_iters = 0;
        return;
    }
    const int n = it->iters;
    it->total_iters += n;
    if (n >= iter_max) it->at_max += 1;
    it->sub_iters += n;
}

int launch_exp_fac_max(qpad_ctx *c, double *out)
{
    ExpFacArgs a;
    memset(&a, 0, sizeof(a));
    a.nr = c->g.nr;
    a.out = (unsigned long long *)c->sc_ef;
    int64_t nmax = 0;
    for (int s = 0; s < c->nspecies; s++) {
        a.gamma[s] = c->sp[s].gamma;
        a.p3[s] = c->sp[s].p3;
        a.cs[s] = c->sp[s].cell_start;
        nmax = std::max(nmax, c->sp[s].npp);
    }
    QPAD_TRY(launch_fill(c, c->sc_ef, 1, 1.0));
    if (c->nspecies > 0 && nmax > 0) {
        const int64_t blocks = std::min<int64_t>((nmax + 255) / 256, 1024);
        hipLaunchKernelGGL(k_exp_fac_max, dim3((unsigned)blocks, (unsigned)c->nspecies), dim3(256), 0, c->stream, a);
        QPAD_HIP_CHECK(hipGetLastError());
    }
    QPAD_HIP_CHECK(hipMemcpyAsync(c->sc_ef_host, c->sc_ef, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    QPAD_HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = *c->sc_ef_host;
    return 0;
}

int launch_iter_fold(qpad_ctx *c, bool reset)
{
    hipLaunchKernelGGL(k_iter_fold, dim3(1), dim3(64), 0, c->stream, c->it, c->g.iter_max, reset ? 1 : 0);
    QPAD_HIP_CHECK(hipGetLastError());
    return 0;
}

// get_subcyc_step (simulation_subcyc_class.f03:431-449)
int subcyc_step(double exp_fac, double exp_fac_max, double dt, double dt_min, double *dt_sub)
{
    int n = 1;
    double d = dt;
    if (exp_fac > exp_fac_max) {
        const double nc = ceil(exp_fac / exp_fac_max);
        d = nc < 2147483647.0 ? dt / nc : 0.0;
        n = nc < 2147483647.0 ? (int)nc : 0;
        if (d < dt_min) {
            n = (int)floor(dt / dt_min);
            d = dt / (double)n;
        }
    }
    *dt_sub = d;
    return n;
}

}  // namespace qpad
