// k_sort.hip -- per-slice re-sort of a plasma species by radial cell
// (the order QPAD's update_bound leaves, part2d_class.f03:2307-2353, is not
// kept: particles are regrouped by cell so that one wave deposits one cell).
//
// Counting sort over the nr + 1 cell keys (key nr = left the box, kept at the
// tail), three launches:
//   k_sort_hist     chunk of kSortChunk particles per workgroup: histogram
//                   of its keys in LDS -> hist[chunk][bin];
//   k_sort_scan     per bin, exclusive prefix of the chunk counts (in place)
//                   and the bin totals;
//   k_sort_scatter  per chunk: bin starts from the totals (workgroup 0 writes
//                   them as cell_start); rank of each particle among the chunk's
//                   particles with the same key, in the fixed order (round,
//                   wave, lane), from wave ballots; position = cell_start +
//                   chunk offset + rank; moves every particle array.
//                   Workgroup 0 also lists the deposit's overflow chunks
//                   (cells holding more than kDepChunk particles).
// Integer counting only, in a fixed order: the permutation is the same on
// every run.
#include <algorithm>
#include <cstring>
#include <string>

#include "kernels.h"

namespace qpad {

constexpr int kSortThreads = 512;   // 8 waves: two per SIMD with one chunk per CU
constexpr int kSortPer = 4;
constexpr int kSortChunk = kSortThreads * kSortPer;
constexpr int kScanGroups = 16;       // chunk groups per bin in k_sort_scan
constexpr int kScanBins = kSortThreads / kScanGroups;

struct SortArgs {
    int64_t n;
    int nb;                 // bins: keys 0 .. nb-1
    int nchunks;
    int kbits;              // bits of the largest key
    const uint32_t *key;
    uint32_t *key_out;
    int *hist;              // [nchunks][nb]
    int *tot;               // [nb]
    int32_t *cs;            // [nb + 1]
    const double *s[8];
    double *d[8];
    const int32_t *sid;
    int32_t *did;
    int32_t *os;            // [nb] overflow rows of cell c: os[c] .. os[c+1]-1
    int4 *desc;             // [n_ov_max] (cell, first particle, end, 0)
    int n_ov_max;
};

// the arguments of every species of one launch (workgroup rows by species)
struct SortMulti {
    SortArgs s[kMaxSpecies];
};

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const SortMulti m)
{
    const SortArgs a = m.s[__builtin_amdgcn_readfirstlane(blockIdx.y)];
    if ((int)blockIdx.x >= a.nchunks) return;   // the species has fewer chunks than the grid
    extern __shared__ int cnt[];
    const int t = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * kSortChunk;
    for (int b = t; b < a.nb; b += kSortThreads) cnt[b] = 0;
    uint32_t kv[kSortPer];
#pragma unroll
    for (int j = 0; j < kSortPer; j++) {
        const int64_t e = e0 + j * kSortThreads + t;
        kv[j] = a.key[min(e, a.n - 1)];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortPer; j++) {
        const int64_t e = e0 + j * kSortThreads + t;
        if (e < a.n) atomicAdd(&cnt[min(kv[j], (uint32_t)(a.nb - 1))], 1);
    }
    __syncthreads();
    int *h = a.hist + (size_t)blockIdx.x * a.nb;
    for (int b = t; b < a.nb; b += kSortThreads) h[b] = cnt[b];
}

// workgroup = kScanBins bins x kScanGroups chunk groups; each thread prefixes
// a contiguous run of at most RUN chunks of its bin (held in registers; RUN
// is the smallest of 8, 16, 32, 64 that covers nchunks / kScanGroups), the
// groups are then prefixed in LDS; the last workgroup turns the bin totals
// into cell_start
constexpr int kScanRun = 64;
constexpr int kMaxSortChunks = kScanGroups * kScanRun;

template <int RUN>
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(const SortMulti m)
{
    const SortArgs a = m.s[__builtin_amdgcn_readfirstlane(blockIdx.y)];
    __shared__ int gsum[kScanGroups][kScanBins];
    const int t = threadIdx.x;
    const int lb = t % kScanBins, grp = t / kScanBins;
    const int b = blockIdx.x * kScanBins + lb;
    const int per = (a.nchunks + kScanGroups - 1) / kScanGroups;   // <= RUN
    const int c0 = grp * per;
    // clamped (always valid) addresses, masked values: no branches between
    // the loads, which all issue together
    const int bc = min(b, a.nb - 1);
    int h[RUN];
#pragma unroll
    for (int u = 0; u < RUN; u++) {
        const int c = min(c0 + u, max(a.nchunks - 1, 0));
        const int x = a.hist[(size_t)c * a.nb + bc];
        h[u] = (b < a.nb && u < per && c0 + u < a.nchunks) ? x : 0;
    }
    int run = 0;
#pragma unroll
    for (int u = 0; u < RUN; u++) {
        const int v = h[u];
        h[u] = run;
        run += v;
    }
    gsum[grp][lb] = run;
    __syncthreads();
    int goff = 0, total = 0;
    for (int q = 0; q < kScanGroups; q++) {
        const int v = gsum[q][lb];
        if (q < grp) goff += v;
        total += v;
    }
#pragma unroll
    for (int u = 0; u < RUN; u++) {
        const int c = c0 + u;
        if (b < a.nb && u < per && c < a.nchunks) a.hist[(size_t)c * a.nb + b] = h[u] + goff;
    }
    if (grp == 0 && b < a.nb) a.tot[b] = total;
}

// exclusive scan of the bin totals into LDS (every scatter workgroup does it:
// no cross-workgroup fence, which would write back the whole L2)
__device__ inline void scan_totals(const SortArgs &a, int *base)
{
    __shared__ int wsum[kSortThreads / 64];
    const int t = threadIdx.x;
    for (int q0 = 0; q0 < a.nb; q0 += 8 * kSortThreads) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = a.tot[min(q0 + u * kSortThreads + t, a.nb - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int q = q0 + u * kSortThreads + t;
            if (q < a.nb) base[q] = v[u];
        }
    }
    __syncthreads();
    const int bper = (a.nb + kSortThreads - 1) / kSortThreads;
    const int b0 = t * bper;
    int local = 0;
    for (int q = 0; q < bper; q++)
        if (b0 + q < a.nb) local += base[b0 + q];
    const int lane = t & 63, wave = t >> 6;
    int incl = local;
    for (int dd = 1; dd < 64; dd <<= 1) {
        const int y = __shfl_up(incl, dd, 64);
        if (lane >= dd) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int run = incl - local;
    for (int w = 0; w < wave; w++) run += wsum[w];
    for (int q = 0; q < bper; q++) {
        if (b0 + q < a.nb) {
            const int v = base[b0 + q];
            base[b0 + q] = run;
            run += v;
        }
    }
    __syncthreads();
}

// workgroup 0, after scan_totals: cell c < nr (nr = nb - 1) with n_c
// particles is deposited in chunks of kDepChunk; chunk 0 by the cell's own
// wave, chunks 1 .. ceil(n_c / kDepChunk) - 1 by overflow waves.  Their rows
// are numbered cell by cell, chunk by chunk: os = exclusive prefix of the
// chunk counts, desc[row] = (cell, first particle, end)
__device__ inline void overflow_chunks(const SortArgs &a, const int *base)
{
    __shared__ int wsum[kSortThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nr = a.nb - 1;
    const int per = (nr + kSortThreads - 1) / kSortThreads;
    const int c0 = t * per;
    int local = 0;
    for (int q = 0; q < per; q++) {
        const int c = c0 + q;
        if (c < nr) {
            const int n = base[c + 1] - base[c];
            local += n > 0 ? (n - 1) / kDepChunk : 0;
        }
    }
    int incl = local;
    for (int dd = 1; dd < 64; dd <<= 1) {
        const int y = __shfl_up(incl, dd, 64);
        if (lane >= dd) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int run = incl - local;
    for (int w = 0; w < wave; w++) run += wsum[w];
    for (int q = 0; q < per; q++) {
        const int c = c0 + q;
        if (c >= nr) break;
        const int p0 = base[c], p1 = base[c + 1];
        const int k = p1 > p0 ? (p1 - p0 - 1) / kDepChunk : 0;
        a.os[c] = min(run, a.n_ov_max);
        for (int u = 1; u <= k; u++) {
            const int row = run + u - 1;
            if (row < a.n_ov_max) a.desc[row] = make_int4(c, p0 + u * kDepChunk, min(p0 + (u + 1) * kDepChunk, p1), 0);
        }
        run += k;
    }
    // the last thread's running count is the total (its range may be empty);
    // rows never exceed n_ov_max (the chunks of n particles number < n / kDepChunk)
    if (t == kSortThreads - 1) a.os[nr] = min(run, a.n_ov_max);
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const SortMulti m)
{
    const SortArgs a = m.s[__builtin_amdgcn_readfirstlane(blockIdx.y)];
    // at least one workgroup per species (workgroup 0 writes cell_start)
    if ((int)blockIdx.x >= max(a.nchunks, 1)) return;
    extern __shared__ int base[];                                  // [nb] bin starts
    unsigned short *cnt = reinterpret_cast<unsigned short *>(base + a.nb);   // [nb] < kSortChunk
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t e0 = (int64_t)blockIdx.x * kSortChunk;
    for (int b = t; b < a.nb; b += kSortThreads) cnt[b] = 0;
    scan_totals(a, base);
    if (blockIdx.x == 0) {
        for (int b = t; b < a.nb; b += kSortThreads) a.cs[b] = base[b];
        if (t == 0) a.cs[a.nb] = (int32_t)a.n;
        overflow_chunks(a, base);
    }
    if (a.n == 0) return;
    uint32_t kv[kSortPer];
#pragma unroll
    for (int j = 0; j < kSortPer; j++) {
        const int64_t e = e0 + j * kSortThreads + t;
        kv[j] = min(a.key[min(e, a.n - 1)], (uint32_t)(a.nb - 1));
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int pos[kSortPer];
#pragma unroll
    for (int j = 0; j < kSortPer; j++) {
        const int64_t e = e0 + j * kSortThreads + t;
        const bool valid = e < a.n;
        const uint32_t key = kv[j];
        uint64_t match = __ballot(valid);
        for (int bit = 0; bit < a.kbits; bit++) {
            const bool bv = (key >> bit) & 1u;
            const uint64_t bal = __ballot(bv);
            match &= bv ? bal : ~bal;
        }
        const int rank = __popcll(match & lt);
        const int count = __popcll(match);
        const bool leader = valid && (match >> lane) == 1ull;
        // waves of this round in order: read the running count, then the
        // group's highest lane advances it
        for (int w = 0; w < kSortThreads / 64; w++) {
            if (wave == w && valid) {
                const int c0 = cnt[key];
                pos[j] = c0 + rank;
                if (leader) cnt[key] = (unsigned short)(c0 + count);
            }
            __syncthreads();
        }
    }
    const int *off = a.hist + (size_t)blockIdx.x * a.nb;
    int64_t dp[kSortPer];
#pragma unroll
    for (int j = 0; j < kSortPer; j++) dp[j] = (int64_t)base[kv[j]] + off[kv[j]] + pos[j];
    // every load of the chunk first (all in flight), then the stores
    double v[8][kSortPer];
    int32_t vi[kSortPer];
#pragma unroll
    for (int j = 0; j < kSortPer; j++) {
        const int64_t e = min(e0 + j * kSortThreads + t, a.n - 1);
#pragma unroll
        for (int f = 0; f < 8; f++) v[f][j] = a.s[f][e];
        vi[j] = a.sid[e];
    }
#pragma unroll
    for (int j = 0; j < kSortPer; j++) {
        const int64_t e = e0 + j * kSortThreads + t;
        if (e >= a.n) continue;
#pragma unroll
        for (int f = 0; f < 8; f++) a.d[f][dp[j]] = v[f][j];
        a.did[dp[j]] = vi[j];
    }
#pragma unroll
    for (int j = 0; j < kSortPer; j++) {
        const int64_t e = e0 + j * kSortThreads + t;
        if (e < a.n) a.key_out[dp[j]] = kv[j];
    }
}

// dynamic LDS of k_sort_scatter: bin starts + counts
static size_t scatter_lds(int nb)
{
    return (sizeof(int) + sizeof(unsigned short)) * (size_t)nb;
}

// allow the scatter's LDS for the largest nr (within the default 64 KB
// limit today; kept explicit so that a larger count type stays legal)
int sort_kernels_init()
{
    QPAD_HIP_CHECK(hipFuncSetAttribute((const void *)k_sort_scatter, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)scatter_lds(8192 + 1)));
    return 0;
}

int sort_workspace(int64_t npmax, int nr, size_t *hist_ints)
{
    const int64_t nchunks = (npmax + kSortChunk - 1) / kSortChunk;
    if (nchunks > kMaxSortChunks) {
        set_error("too many particles per species for the cell sort (max " +
                  std::to_string((long long)kMaxSortChunks * kSortChunk) + ")");
        return QPAD_E_INVALID;
    }
    *hist_ints = (size_t)nchunks * (size_t)(nr + 1);
    return 0;
}

static SortArgs sort_args(qpad_ctx *c, int s)
{
    Species &sp = c->sp[s];
    const int64_t n = sp.npp;
    SortArgs a;
    memset(&a, 0, sizeof(a));
    a.n = n;
    a.nb = c->g.nr + 1;
    a.nchunks = (int)((n + kSortChunk - 1) / kSortChunk);
    int kb = 1;
    while ((1 << kb) <= a.nb - 1) kb++;
    a.kbits = kb;
    a.key = sp.key;
    a.key_out = sp.key_alt;
    a.hist = sp.hist;
    a.tot = sp.tot;
    a.cs = sp.cell_start;
    const double *src[8] = {sp.x1, sp.x2, sp.p1, sp.p2, sp.p3, sp.q, sp.gamma, sp.psi};
    double *dst[8] = {sp.a_x1, sp.a_x2, sp.a_p1, sp.a_p2, sp.a_p3, sp.a_q, sp.a_gamma, sp.a_psi};
    for (int f = 0; f < 8; f++) {
        a.s[f] = src[f];
        a.d[f] = dst[f];
    }
    a.sid = sp.id;
    a.did = sp.a_id;
    a.os = sp.ov_start;
    a.desc = sp.ov_desc;
    a.n_ov_max = sp.n_ov_max;
    return a;
}

// species s0 .. s1-1 sorted by the same three launches (workgroup rows by
// species; each species' rows beyond its own chunk count return at once)
static int launch_sort_range(qpad_ctx *c, int s0, int s1)
{
    SortMulti m;
    memset(&m, 0, sizeof(m));
    int maxc = 0;
    for (int s = s0; s < s1; s++) {
        m.s[s - s0] = sort_args(c, s);
        maxc = std::max(maxc, m.s[s - s0].nchunks);
    }
    const int ns = s1 - s0, nb = c->g.nr + 1;
    const size_t lds = sizeof(int) * (size_t)nb;
    if (maxc > 0) hipLaunchKernelGGL(k_sort_hist, dim3(maxc, ns), dim3(kSortThreads), lds, c->stream, m);
    {
        const dim3 sg((nb + kScanBins - 1) / kScanBins, ns), sb(kSortThreads);
        const int per = (maxc + kScanGroups - 1) / kScanGroups;
        if (per <= 8) hipLaunchKernelGGL(k_sort_scan<8>, sg, sb, 0, c->stream, m);
        else if (per <= 16) hipLaunchKernelGGL(k_sort_scan<16>, sg, sb, 0, c->stream, m);
        else if (per <= 32) hipLaunchKernelGGL(k_sort_scan<32>, sg, sb, 0, c->stream, m);
        else hipLaunchKernelGGL(k_sort_scan<kScanRun>, sg, sb, 0, c->stream, m);
    }
    // at least one scatter workgroup per species: workgroup 0 also writes cell_start
    hipLaunchKernelGGL(k_sort_scatter, dim3(std::max(maxc, 1), ns), dim3(kSortThreads), scatter_lds(nb), c->stream,
                       m);
    QPAD_HIP_CHECK(hipGetLastError());
    for (int s = s0; s < s1; s++) {
        Species &sp = c->sp[s];
        if (sp.npp > 0) {
            std::swap(sp.x1, sp.a_x1); std::swap(sp.x2, sp.a_x2);
            std::swap(sp.p1, sp.a_p1); std::swap(sp.p2, sp.a_p2); std::swap(sp.p3, sp.a_p3);
            std::swap(sp.q, sp.a_q); std::swap(sp.gamma, sp.a_gamma); std::swap(sp.psi, sp.a_psi);
            std::swap(sp.id, sp.a_id);
            std::swap(sp.key, sp.key_alt);
        }
    }
    return 0;
}

int launch_sort(qpad_ctx *c, int s) { return launch_sort_range(c, s, s + 1); }

int launch_sort_all(qpad_ctx *c) { return c->nspecies ? launch_sort_range(c, 0, c->nspecies) : 0; }

}  // namespace qpad
