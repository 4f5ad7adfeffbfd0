// Non-templated kernels: delay-ring flush, growth, history update and index-map construction.
// Included by bghip.hip only.
#pragma once
#include "bghip_kernels.h"
#include <cfloat>

// Inserts the spikes of the last fused block into the delay rings so that between API calls the
// rings hold every pending spike, as the reference's rings do after each step.
__global__ __launch_bounds__(BG_BLOCK) void k_flush_block(DevNet n, uint64_t tprev0, int sel)
{
    uint64_t slot = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (slot >= n.S) return;
    const uint32_t meta = n.meta[slot];
    if (!(meta & BG_META_IN_USE)) return;
    uint32_t fm = n.mask[sel][meta & BG_META_SRC_MASK];
    if (!fm) return;
    const uint32_t delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
    const uint32_t Rm = (uint32_t)n.R - 1;
    while (fm) {
        const int b0 = __ffs(fm) - 1;
        fm &= fm - 1;
        const uint32_t b = (uint32_t)((tprev0 + (uint64_t)b0 + delay) & Rm);
        uint32_t *w = &n.ring[(uint64_t)(b >> 5) * n.S + slot];
        const uint32_t m = 1u << (b & 31);
        if (*w & m) dev_error(n, 1u);
        *w |= m;
    }
}

// ---------------------------------------------------------------------------------------------
// Growth. ConnGrowth::updateConns (Connections/ConnGrowth.cpp:295-311): rates, outgrowth, radii.
__global__ void k_growth_radii(DevNet n, float epochDuration, float epsilon, float maxRate, float beta, float c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n.N) return;
    if (!(n.cnt[i] < n.max_spikes)) dev_error(n, 2u);
    float rate = (float)n.cnt[i] / epochDuration;
    n.rates[i] = rate;
    float x = epsilon - rate / maxRate;
    x = x / beta;
    x = bg_expf(x);
    x = x + 1.0f;
    x = 2.0f / x;
    float outgrowth = 1.0f - x;
    n.radii[i] = n.radii[i] + outgrowth * c;
}

struct CreateCtx {
    float W_adj;            // SYNAPSE_STRENGTH_ADJUSTMENT (AllSynapses.h:190)
    int32_t total_delay[4];
    float U[4], D[4], F[4];
    float u0, r0;
    float Apos, Aneg, gap, tauspost, tauspre, taupos, tauneg, Wex, mupos, muneg;
    int32_t useFD;
    int syn_kind;
    uint64_t step;          // creation step (delay-ring phase)
    int coh;                // cohort of the created synapses (implicit-ring networks)
};

__device__ __forceinline__ int syn_type_of(int ts, int td)
{   // Layouts/Layout.cpp:129-141 (INH = 1, EXC = 2)
    if (ts == 1) return td == 1 ? 0 : 1;
    return td == 1 ? 2 : 3;
}
__device__ __forceinline__ int syn_sign_of(int type) { return type <= 1 ? -1 : 1; }

// createSynapse (Synapses/AllSpikingSynapses.cpp:191-233 and subclasses) for slot `slot`.
__device__ void create_synapse(const DevNet &n, const CreateCtx &c, uint64_t slot, int src, int type, float w)
{
    n.meta[slot] = (uint32_t)src | ((uint32_t)type << BG_META_TYPE_SHIFT) |
                   ((uint32_t)c.total_delay[type] << BG_META_DELAY_SHIFT) |
                   ((uint32_t)(c.coh & (BG_COH_MAX - 1)) << BG_META_COH_SHIFT) | BG_META_IN_USE;
    n.phase[slot] = (uint8_t)(c.step & (uint64_t)(n.R - 1));
    n.psr[slot] = 0.0f;
    n.W[slot] = w;
    if (n.Wd) n.Wd[slot] = w / n.decay[type];
    if (c.syn_kind == SK_STDP || c.syn_kind == SK_DYNSTDP) {   // byte planes (sring_word)
        for (int q = 0; q < 4 * n.W32; q++) ((uint8_t *)n.ring)[(uint64_t)q * n.S + slot] = 0u;
    } else {
        for (int q = 0; q < n.W32; q++) n.ring[(uint64_t)q * n.S + slot] = 0u;
    }
    if (c.syn_kind == SK_DS || c.syn_kind == SK_DYNSTDP) {
        n.r[slot] = c.r0;
        n.us[slot] = c.u0;
        n.last[slot] = BG_NO_SPIKE;
        n.U[slot] = c.U[type];
        n.Dd[slot] = c.D[type];
        n.F[slot] = c.F[type];
    }
    if (c.syn_kind == SK_STDP || c.syn_kind == SK_DYNSTDP) {
        n.Apos[slot] = c.Apos; n.Aneg[slot] = c.Aneg; n.gap[slot] = c.gap; n.tauspost[slot] = c.tauspost;
        n.tauspre[slot] = c.tauspre; n.taupos[slot] = c.taupos; n.tauneg[slot] = c.tauneg; n.Wex[slot] = c.Wex;
        n.mupos[slot] = c.mupos; n.muneg[slot] = c.muneg; n.useFD[slot] = (uint8_t)c.useFD;
    }
}

#define BG_GROWTH_TILE 1024
#define BG_CONN_WORDS (BG_MAX_NEURONS / 32)
// ConnGrowth::updateSynapsesWeights (Connections/ConnGrowth.cpp:388-463), one workgroup per
// destination neuron (grid-strided). Existing synapses get W = area * sign * 1e-8 or are erased
// when area < 0; every unconnected source with area > 0 gets a synapse in dest's first free slot,
// in ascending source order. Source coordinates and radii stream through LDS in tiles; the
// connected-source set is an LDS bitmap; create decisions are compacted to slots with a
// wavefront-ballot + workgroup prefix scan.
// W(s, d) of the update: the circles' overlap area (the growth rule's W = area,
// Connections/ConnGrowth.cpp:395), or, when Wt is set, the caller's dense matrix W(s, d) held
// transposed (Wt[d * N + s]), so that a destination's workgroup reads its sources contiguously.
__device__ __forceinline__ float pair_weight(const DevNet &n, const float *Wt, int s, int d, float xs, float ys,
                                             float rs, float xd, float yd, float rd)
{
    return Wt ? Wt[(uint64_t)d * (uint64_t)n.N + (uint64_t)s] : overlap_area(xs, ys, rs, xd, yd, rd, s == d);
}

// Bounding box and largest radius of each BG_GROWTH_TILE-source tile, for k_growth_dest to pass over
// tiles whose sources cannot overlap a destination (one workgroup per tile; min and max are exact
// in any order).
__global__ __launch_bounds__(BG_BLOCK) void k_growth_tiles(DevNet n)
{
    __shared__ float red[5][BG_BLOCK / BG_WAVE];
    const int t = blockIdx.x, tid = threadIdx.x, lane = tid & (BG_WAVE - 1), wv = tid / BG_WAVE;
    const int s0 = t * BG_GROWTH_TILE, len = min(BG_GROWTH_TILE, n.N - s0);
    float v[5] = {FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int q = tid; q < len; q += BG_BLOCK) {
        const float x = n.xloc[s0 + q], y = n.yloc[s0 + q], r = n.radii[s0 + q];
        v[0] = fminf(v[0], x); v[1] = fmaxf(v[1], x);
        v[2] = fminf(v[2], y); v[3] = fmaxf(v[3], y);
        v[4] = fmaxf(v[4], r);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v[0] = fminf(v[0], __shfl_xor(v[0], off, 64)); v[1] = fmaxf(v[1], __shfl_xor(v[1], off, 64));
        v[2] = fminf(v[2], __shfl_xor(v[2], off, 64)); v[3] = fmaxf(v[3], __shfl_xor(v[3], off, 64));
        v[4] = fmaxf(v[4], __shfl_xor(v[4], off, 64));
    }
    if (lane == 0)
        for (int k = 0; k < 5; k++) red[k][wv] = v[k];
    __syncthreads();
    if (tid < 5) {
        float a = red[tid][0];
        for (int w = 1; w < BG_BLOCK / BG_WAVE; w++) a = (tid & 1) || tid == 4 ? fmaxf(a, red[tid][w]) : fminf(a, red[tid][w]);
        n.gbox[5 * t + tid] = a;
    }
}

// True if no source of the tile with box b can overlap destination (xd, yd, rd): overlap_area is
// zero unless the distance is below the sum of the radii, and every source lies at least the box
// distance away with at most the tile's largest radius. The margin far exceeds the rounding of
// overlap_area's float distance and sum, so a passed-over tile holds no pair with a positive area.
__device__ __forceinline__ bool growth_tile_far(const float *b, float xd, float yd, float rd)
{
    const float dx = fmaxf(0.0f, fmaxf(b[0] - xd, xd - b[1]));
    const float dy = fmaxf(0.0f, fmaxf(b[2] - yd, yd - b[3]));
    const float reach = (rd + b[4]) * 1.0001f;
    return dx * dx + dy * dy > reach * reach;
}

// In-place transpose of a square N x N float matrix, one 32 x 32 tile pair per workgroup
// (tile (i, j) with i <= j swaps with tile (j, i) through LDS). Grid: (T, T), T = ceil(N / 32),
// 256 threads.
#define BG_TT 32
__global__ __launch_bounds__(BG_BLOCK) void k_transpose_square(float *W, int N)
{
    __shared__ float a[BG_TT][BG_TT + 1], b[BG_TT][BG_TT + 1];
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (ti > tj) return;
    const int tx = threadIdx.x & (BG_TT - 1), ty = threadIdx.x / BG_TT;
    const int rows = BG_BLOCK / BG_TT;
    for (int r = ty; r < BG_TT; r += rows) {
        const int i = ti * BG_TT + r, j = tj * BG_TT + tx;
        if (i < N && j < N) a[r][tx] = W[(size_t)i * N + j];
        const int i2 = tj * BG_TT + r, j2 = ti * BG_TT + tx;
        if (ti != tj && i2 < N && j2 < N) b[r][tx] = W[(size_t)i2 * N + j2];
    }
    __syncthreads();
    for (int r = ty; r < BG_TT; r += rows) {
        // element (i, j) of the transpose is W(j, i): tile (ti, tj) entry (tx, r)
        const int i = tj * BG_TT + r, j = ti * BG_TT + tx;
        if (i < N && j < N) W[(size_t)i * N + j] = a[tx][r];
        const int i2 = ti * BG_TT + r, j2 = tj * BG_TT + tx;
        if (ti != tj && i2 < N && j2 < N) W[(size_t)i2 * N + j2] = b[tx][r];
    }
}

__device__ void growth_serial(const DevNet &n, const CreateCtx &c, int d, const uint32_t *conn, uint32_t *stats,
                              const float *Wd)
{
    // Replays the reference's per-source order one source at a time on one lane. Used only for rows
    // with holes or erasures whose capacity M exceeds the workgroup path's slot bitmaps
    // (growth_holed below).
    const int N = n.N, M = n.M;
    const uint64_t base = (uint64_t)d * M;
    const float xd = n.xloc[d], yd = n.yloc[d], rd = n.radii[d];
    const int td = n.ntype[d];
    uint32_t count = n.syn_count[d], hi = n.slot_hi[d], created = 0, erased = 0;
    for (int s = 0; s < N; s++) {
        float area = pair_weight(n, Wd, s, d, n.xloc[s], n.yloc[s], n.radii[s], xd, yd, rd);
        int type = syn_type_of(n.ntype[s], td);
        bool connected = (conn[s >> 5] >> (s & 31)) & 1u;
        if (connected) {
            if (area < 0) {
                for (uint32_t k = 0; k < hi; k++) {
                    uint32_t meta = n.meta[base + k];
                    if ((meta & BG_META_IN_USE) && (int)(meta & BG_META_SRC_MASK) == s) {
                        n.meta[base + k] = 0u;
                        count--;
                        erased++;
                    }
                }
            }
        } else if (area > 0) {
            if (count >= (uint32_t)M) { dev_error(n, 4u); continue; }
            uint32_t k = 0;
            while (n.meta[base + k] & BG_META_IN_USE) k++;
            create_synapse(n, c, base + k, s, type, area * (float)syn_sign_of(type) * c.W_adj);
            if (k + 1 > hi) hi = k + 1;
            count++;
            created++;
        }
    }
    while (hi > 0 && !(n.meta[base + hi - 1] & BG_META_IN_USE)) hi--;
    n.syn_count[d] = count;
    n.slot_hi[d] = hi;
    if (!(count < (uint32_t)M)) dev_error(n, 4u);
    atomicAdd(&stats[0], created);
    atomicAdd(&stats[1], erased);
}

// Rows with holes or erasures, whole workgroup: the reference visits the sources in ascending
// order and, for destination d, erases a connected source's synapses when W < 0 and gives an
// unconnected source with W > 0 a synapse in d's first free slot (Connections/ConnGrowth.cpp:
// 388-463, AllSynapses::addSynapse at Synapses/AllSynapses.cpp:357-377), so an erasure frees a slot
// that a later source may take. The workgroup evaluates W for 256 sources at a time; wavefront 0
// then applies that chunk's erase and create events in source order, with the row's in-use slots
// and the slots created in this update as LDS bitmaps: a create takes the lowest clear in-use bit
// (a ballot over 64 bitmap words per round), an erase finds the source's slots with a ballot over
// 64 slots per round. Same decisions, slots and order as the one-lane replay growth_serial.
#define BG_HOLED_MAX_M 8192
#define BG_HOLED_WORDS (BG_HOLED_MAX_M / 32)

// LDS written by some lanes of a wavefront and read by others afterwards
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct HoledLds {
    uint32_t inuse[BG_HOLED_WORDS];   // slot k of the row in use
    uint32_t made[BG_HOLED_WORDS];    // slot k created by this update (its source is unconnected)
    float area[BG_BLOCK];             // W(s0 + t, d) of the current chunk
    uint64_t cre[BG_BLOCK / BG_WAVE], era[BG_BLOCK / BG_WAVE];
};

__device__ void growth_holed(const DevNet &n, const CreateCtx &c, int d, const uint32_t *conn, uint32_t *stats,
                             const float *Wt, HoledLds &L, float xd, float yd, float rd, int td, uint32_t cnt0,
                             uint32_t hi0)
{
    const int tid = threadIdx.x, lane = tid & (BG_WAVE - 1), wv = tid / BG_WAVE;
    const int N = n.N, M = n.M;
    const uint64_t base = (uint64_t)d * M;
    const uint32_t words = (uint32_t)(M + 31) >> 5;
    for (uint32_t w = tid; w < words; w += BG_BLOCK) {
        L.inuse[w] = 0u;
        L.made[w] = 0u;
    }
    __syncthreads();
    for (uint32_t k0 = 0; k0 < hi0; k0 += BG_BLOCK) {
        const uint32_t k = k0 + tid;
        const uint64_t b = __ballot(k < hi0 && (n.meta[base + k] & BG_META_IN_USE));
        if (lane == 0 && b) {
            const uint32_t w0 = (k0 + BG_WAVE * wv) >> 5;
            L.inuse[w0] = (uint32_t)b;
            L.inuse[w0 + 1] = (uint32_t)(b >> 32);
        }
    }
    __syncthreads();
    // wavefront 0's running state (uniform across its lanes)
    uint32_t count = cnt0, hi = hi0, created = 0, erased = 0, fw = 0;
    const uint32_t last_mask = (M & 31) ? ((1u << (M & 31)) - 1u) : 0xFFFFFFFFu;
    for (int s0 = 0; s0 < N; s0 += BG_BLOCK) {
        const int s = s0 + tid;
        bool cr = false, er = false;
        float area = 0.0f;
        if (s < N) {
            const bool connected = (conn[s >> 5] >> (s & 31)) & 1u;
            area = pair_weight(n, Wt, s, d, n.xloc[s], n.yloc[s], n.radii[s], xd, yd, rd);
            er = connected && area < 0;
            cr = !connected && area > 0;
        }
        L.area[tid] = area;
        const uint64_t bc = __ballot(cr), be = __ballot(er);
        if (lane == 0) {
            L.cre[wv] = bc;
            L.era[wv] = be;
        }
        __syncthreads();
        if (wv == 0) {
            for (int w = 0; w < BG_BLOCK / BG_WAVE; w++) {
                const uint64_t ew = L.era[w];
                uint64_t ev = L.cre[w] | ew;
                while (ev) {
                    const int j = __ffsll((unsigned long long)ev) - 1;
                    ev &= ev - 1;
                    const int se = s0 + BG_WAVE * w + j;
                    if ((ew >> j) & 1ull) {
                        // erase: every in-use slot of source se that existed before this update
                        for (uint32_t k0 = 0; k0 < hi; k0 += BG_WAVE) {
                            const uint32_t k = k0 + lane;
                            bool hit = false;
                            if (k < hi) {
                                const uint32_t wd = L.inuse[k >> 5] & ~L.made[k >> 5];
                                if ((wd >> (k & 31)) & 1u) hit = (int)(n.meta[base + k] & BG_META_SRC_MASK) == se;
                            }
                            if (hit) n.meta[base + k] = 0u;
                            const uint64_t hb = __ballot(hit);
                            if (hb) {
                                if (lane == 0) {
                                    L.inuse[k0 >> 5] &= ~(uint32_t)hb;
                                    if ((k0 >> 5) + 1 < words) L.inuse[(k0 >> 5) + 1] &= ~(uint32_t)(hb >> 32);
                                }
                                const uint32_t nh = (uint32_t)__popcll(hb);
                                count -= nh;
                                erased += nh;
                                fw = min(fw, k0 >> 5);
                                wave_lds_sync();
                            }
                        }
                    } else {
                        if (count >= (uint32_t)M) {
                            if (lane == 0) dev_error(n, 4u);
                            continue;
                        }
                        // create: the lowest free slot of the row
                        uint32_t k = 0xFFFFFFFFu;
                        for (uint32_t w0 = fw; w0 < words; w0 += BG_WAVE) {
                            const uint32_t ww = w0 + lane;
                            uint32_t fr = 0u;
                            if (ww < words) fr = ~L.inuse[ww] & (ww == words - 1 ? last_mask : 0xFFFFFFFFu);
                            const uint64_t bw = __ballot(fr != 0u);
                            if (bw) {
                                const int l = __ffsll((unsigned long long)bw) - 1;
                                const uint32_t frl = (uint32_t)__shfl((int)fr, l);
                                k = 32u * (w0 + (uint32_t)l) + (uint32_t)(__ffs((int)frl) - 1);
                                break;
                            }
                        }
                        if (k == 0xFFFFFFFFu) {   // unreachable while count < M
                            if (lane == 0) dev_error(n, 4u);
                            continue;
                        }
                        if (lane == 0) {
                            L.inuse[k >> 5] |= 1u << (k & 31);
                            L.made[k >> 5] |= 1u << (k & 31);
                            const int type = syn_type_of(n.ntype[se], td);
                            create_synapse(n, c, base + k, se, type,
                                           L.area[BG_WAVE * w + j] * (float)syn_sign_of(type) * c.W_adj);
                        }
                        wave_lds_sync();
                        fw = k >> 5;
                        hi = max(hi, k + 1);
                        count++;
                        created++;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        uint32_t h = 0;
        for (int w = (int)words - 1; w >= 0; w--)
            if (L.inuse[w]) {
                h = 32u * (uint32_t)w + 32u - (uint32_t)__clz((int)L.inuse[w]);
                break;
            }
        n.syn_count[d] = count;
        n.slot_hi[d] = h;
        if (!(count < (uint32_t)M)) dev_error(n, 4u);
        atomicAdd(&stats[0], created);
        atomicAdd(&stats[1], erased);
    }
}

__global__ __launch_bounds__(BG_BLOCK) void k_growth_dest(DevNet n, CreateCtx c, uint32_t *stats, const float *Wd)
{
    __shared__ float sx[BG_GROWTH_TILE], sy[BG_GROWTH_TILE], sr[BG_GROWTH_TILE];
    __shared__ uint32_t conn[BG_CONN_WORDS];
    __shared__ uint32_t s_wave[BG_BLOCK / BG_WAVE];
    __shared__ uint32_t s_base, s_erase;
    __shared__ HoledLds holed;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int N = n.N, M = n.M;
    const int nwords = (N + 31) >> 5;
    for (int d = blockIdx.x; d < N; d += gridDim.x) {
        const float xd = n.xloc[d], yd = n.yloc[d], rd = n.radii[d];
        const int td = n.ntype[d];
        const uint64_t base = (uint64_t)d * M;
        const uint32_t cnt0 = n.syn_count[d], hi0 = n.slot_hi[d];
        __syncthreads();
        for (int w = tid; w < nwords; w += BG_BLOCK) conn[w] = 0u;
        if (tid == 0) { s_erase = 0; s_base = 0; }
        __syncthreads();
        // existing synapses: adjust W or flag an erase
        for (uint32_t k = tid; k < hi0; k += BG_BLOCK) {
            uint32_t meta = n.meta[base + k];
            if (!(meta & BG_META_IN_USE)) continue;
            int s = (int)(meta & BG_META_SRC_MASK);
            atomicOr(&conn[s >> 5], 1u << (s & 31));
            float area = pair_weight(n, Wd, s, d, n.xloc[s], n.yloc[s], n.radii[s], xd, yd, rd);
            int type = syn_type_of(n.ntype[s], td);
            if (area < 0) {
                atomicAdd(&s_erase, 1u);
            } else {
                const float w = area * (float)syn_sign_of(type) * c.W_adj;
                n.W[base + k] = w;
                if (n.Wd) n.Wd[base + k] = w / n.decay[type];
            }
        }
        __syncthreads();
        if (s_erase || cnt0 != hi0) {
            if (M <= BG_HOLED_MAX_M) growth_holed(n, c, d, conn, stats, Wd, holed, xd, yd, rd, td, cnt0, hi0);
            else if (tid == 0) growth_serial(n, c, d, conn, stats, Wd);
            continue;
        }
        // rows without holes: the j-th creation (ascending source) takes slot hi0 + j; with overlap
        // weights, tiles no source of which reaches d create nothing and are passed over
        for (int s0 = 0; s0 < N; s0 += BG_GROWTH_TILE) {
            if (!Wd && growth_tile_far(n.gbox + 5 * (s0 / BG_GROWTH_TILE), xd, yd, rd)) continue;
            const int len = min(BG_GROWTH_TILE, N - s0);
            __syncthreads();
            for (int q = tid; q < len; q += BG_BLOCK) {
                sx[q] = n.xloc[s0 + q];
                sy[q] = n.yloc[s0 + q];
                sr[q] = n.radii[s0 + q];
            }
            __syncthreads();
            for (int q0 = 0; q0 < len; q0 += BG_BLOCK) {
                const int q = q0 + tid;
                const int s = s0 + q;
                bool create = false;
                float area = 0.0f;
                if (q < len && !((conn[s >> 5] >> (s & 31)) & 1u)) {
                    area = pair_weight(n, Wd, s, d, sx[q], sy[q], sr[q], xd, yd, rd);
                    create = area > 0;
                }
                const uint64_t ball = __ballot(create);
                const uint32_t lane_rank = (uint32_t)__popcll(ball & ((1ull << lane) - 1ull));
                if (lane == 0) s_wave[wv] = (uint32_t)__popcll(ball);
                __syncthreads();
                uint32_t before = s_base, total = 0;
                for (int w = 0; w < BG_BLOCK / BG_WAVE; w++) {
                    if (w < wv) before += s_wave[w];
                    total += s_wave[w];
                }
                __syncthreads();
                if (create) {
                    const uint32_t rank = before + lane_rank;
                    if (cnt0 + rank >= (uint32_t)M) {
                        dev_error(n, 4u);
                    } else {
                        const int type = syn_type_of(n.ntype[s], td);
                        create_synapse(n, c, base + hi0 + rank, s, type, area * (float)syn_sign_of(type) * c.W_adj);
                    }
                }
                if (tid == 0) s_base += total;
            }
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t created = s_base;
            uint32_t count = cnt0 + created;
            if (count > (uint32_t)M) count = M;
            n.syn_count[d] = count;
            n.slot_hi[d] = count;
            if (!(count < (uint32_t)M)) dev_error(n, 4u);
            atomicAdd(&stats[0], created);
        }
    }
}

// AllSynapses::addSynapse (Synapses/AllSynapses.cpp:357-377) for a list of new synapses grouped by
// destination (entries of destination dests[q] are [begin[q], begin[q+1]) in call order), one
// wavefront per destination. The reference puts each new synapse into the first free slot left
// by the earlier ones, so entry j of a destination lands in its j-th free slot: the wave walks the
// slots 64 at a time, ballots the free ones, and lane L of a chunk takes entry
// (entries placed so far) + (free slots below L). Entries beyond the free slots are dropped with
// the slot-overflow error bit.
// Optional per-entry overrides of createSynapse's per-type values: tdel (total delay in steps) and
// wex (the STDP weight bound Wex); null keeps the per-type value.
__global__ __launch_bounds__(BG_BLOCK) void k_add_synapses(DevNet n, CreateCtx c, const int32_t *dests,
                                                          const int64_t *begin, const int32_t *src,
                                                          const float *w, const int32_t *tdel, const float *wex,
                                                          int nd)
{
    const int lane = threadIdx.x & 63;
    const int q = (int)((blockIdx.x * (uint64_t)BG_BLOCK + threadIdx.x) >> 6);
    if (q >= nd) return;
    const int d = dests[q];
    const int64_t b = begin[q], cnt = begin[q + 1] - b;
    const int M = n.M, td = n.ntype[d];
    const uint64_t base = (uint64_t)d * M;
    int64_t placed = 0;
    uint32_t top = 0;   // one past the highest slot this lane filled
    for (int k0 = 0; k0 < M && placed < cnt; k0 += 64) {
        const int k = k0 + lane;
        const bool free_slot = k < M && !(n.meta[base + k] & BG_META_IN_USE);
        const uint64_t ball = __ballot(free_slot);
        const int64_t j = placed + (int64_t)__popcll(ball & ((1ull << lane) - 1ull));
        if (free_slot && j < cnt) {
            const int s = src[b + j];
            create_synapse(n, c, base + k, s, syn_type_of(n.ntype[s], td), w[b + j]);
            if (tdel)
                n.meta[base + k] = (n.meta[base + k] & ~(0xFFu << BG_META_DELAY_SHIFT)) |
                                   ((uint32_t)tdel[b + j] << BG_META_DELAY_SHIFT);
            if (wex && (c.syn_kind == SK_STDP || c.syn_kind == SK_DYNSTDP)) n.Wex[base + k] = wex[b + j];
            top = (uint32_t)k + 1u;
        }
        placed += (int64_t)__popcll(ball);
    }
    for (int off = 32; off > 0; off >>= 1) top = max(top, (uint32_t)__shfl_xor((int)top, off));
    if (lane == 0) {
        const int64_t made = placed < cnt ? placed : cnt;
        if (made < cnt) dev_error(n, 4u);
        n.syn_count[d] += (uint32_t)made;
        if (top > n.slot_hi[d]) n.slot_hi[d] = top;
    }
}

// Recorder clamp + clearSpikeCounts (Recorders/XmlGrowthRecorder.cpp:89-90,
// Neurons/AllSpikingNeurons.cpp:79-87).
// End of epoch: XmlGrowthRecorder's rates/radii rows with the minRadius clamp
// (Recorders/XmlGrowthRecorder.cpp:82-96; rows only when the recorder keeps them on the device)
// and AllSpikingNeurons::clearSpikeCounts (Neurons/AllSpikingNeurons.cpp:79-87).
__global__ void k_update_history(DevNet n, float minRadius, int clamp, float *rates_row, float *radii_row)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n.N) return;
    if (rates_row) rates_row[i] = n.rates[i];
    if (clamp && n.radii[i] < minRadius) n.radii[i] = minRadius;
    if (radii_row) radii_row[i] = n.radii[i];
    n.off[i] = (n.cnt[i] + n.off[i]) % n.max_spikes;
    n.cnt[i] = 0;
}

// XmlRecorder::compileHistories (Recorders/XmlRecorder.cpp:71-104) on the device: every spike
// recorded since the last clear is counted in 1 s bin int(s*dt) and 10 ms bin int(s*dt*100).
// Integer counts make the result independent of the order of the atomic adds; the reference's
// float bins equal min(count, 2^24) (1.0 added to 2^24 in float rounds back to 2^24), which is
// how counts are converted at download. Bins [b_lo, b_lo + nb_loc) and [s_lo, s_lo + ns_loc)
// (the epoch's step range) are counted in LDS first, one global atomic per block and bin.
__global__ __launch_bounds__(BG_BLOCK) void k_record_spikes(DevNet n, uint32_t *burst, int64_t nb, uint32_t *shist,
                                                            int64_t nsh, int64_t b_lo, int nb_loc, int64_t s_lo,
                                                            int ns_loc)
{
    extern __shared__ uint32_t lds_bins[];
    for (int k = threadIdx.x; k < nb_loc + ns_loc; k += blockDim.x) lds_bins[k] = 0u;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n.N) {
        const int ms = n.max_spikes, cnt = n.cnt[i];
        int idx = n.off[i];
        for (int q = 0; q < cnt; q++, idx++) {
            if (idx >= ms) idx = 0;
            const double x = (double)n.hist[(uint64_t)i * ms + idx] * (double)n.dt;
            const int i1 = (int)x, i2 = (int)(x * 100);
            if (i1 >= 0 && i1 < nb) {
                const int64_t r = (int64_t)i1 - b_lo;
                if (r >= 0 && r < nb_loc) atomicAdd(&lds_bins[r], 1u);
                else atomicAdd(&burst[i1], 1u);
            }
            if (i2 >= 0 && i2 < nsh) {
                const int64_t r = (int64_t)i2 - s_lo;
                if (r >= 0 && r < ns_loc) atomicAdd(&lds_bins[nb_loc + r], 1u);
                else atomicAdd(&shist[i2], 1u);
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nb_loc; k += blockDim.x)
        if (lds_bins[k]) atomicAdd(&burst[b_lo + k], lds_bins[k]);
    for (int k = threadIdx.x; k < ns_loc; k += blockDim.x)
        if (lds_bins[nb_loc + k]) atomicAdd(&shist[s_lo + k], lds_bins[nb_loc + k]);
}

// ---------------------------------------------------------------------------------------------
// Index maps (AllSynapses::createSynapseImap, Synapses/AllSynapses.cpp:223-293).
// Exclusive scan of syn_count into in_begin, single workgroup (N <= 262144); total into *total.
__global__ __launch_bounds__(1024) void k_scan_counts(const uint32_t *counts, uint32_t *begin, int N, uint64_t *total)
{
    __shared__ uint32_t s_part[1024];
    __shared__ uint32_t s_carry;
    const int tid = threadIdx.x;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < N; c0 += 1024) {
        const int i = c0 + tid;
        const uint32_t v = i < N ? counts[i] : 0u;
        s_part[tid] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t x = tid >= off ? s_part[tid - off] : 0u;
            __syncthreads();
            s_part[tid] += x;
            __syncthreads();
        }
        if (i < N) begin[i] = s_carry + s_part[tid] - v;
        __syncthreads();
        if (tid == 1023) s_carry += s_part[1023];
        __syncthreads();
    }
    if (tid == 0) *total = s_carry;
}

// Destinations ordered by row length, longest first (a counting sort over lengths capped at 1023,
// one workgroup): the implicit-ring fused kernels hand out their workgroups in this order, so the
// longest rows start first and the short ones fill the end of a launch. Within a length the order
// is whatever the atomics give; no result depends on it.
__global__ __launch_bounds__(1024) void k_lpt_order(const uint32_t *counts, uint32_t *order, int N)
{
    __shared__ uint32_t s_hist[1024];
    const int tid = threadIdx.x;
    s_hist[tid] = 0;
    __syncthreads();
    for (int i = tid; i < N; i += 1024) atomicAdd(&s_hist[1023u - min(counts[i], 1023u)], 1u);
    __syncthreads();
    if (tid == 0) {
        uint32_t run = 0;
        for (int b = 0; b < 1024; b++) {
            const uint32_t c = s_hist[b];
            s_hist[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int i = tid; i < N; i += 1024) order[atomicAdd(&s_hist[1023u - min(counts[i], 1023u)], 1u)] = (uint32_t)i;
}

// The per-step STDP streaming pass reads a synapse's type from this byte array instead of its 4-byte
// meta (meta is re-read only for rows whose destination fired): one thread per slot.
__global__ void k_slot_types(DevNet n)
{
    const uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s < n.S) n.stype[s] = (uint8_t)((n.meta[s] >> BG_META_TYPE_SHIFT) & 3u);
}

// The per-step STDP neuron pass pushes a spike to each synapse of the outgoing map at its delay;
// keeping that delay beside the map entry saves the push a dependent read of meta. One thread per
// entry, after the map is sorted.
__global__ void k_out_delays(DevNet n, uint64_t total)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < total) n.out_delay[i] = (uint8_t)((n.meta[n.out_map[i]] >> BG_META_DELAY_SHIFT) & 0xFFu);
}

// incoming map: in-use slots of each destination in ascending slot order (wave per dest).
__global__ __launch_bounds__(BG_BLOCK) void k_fill_incoming(DevNet n)
{
    const int lane = threadIdx.x & 63;
    const int d = (int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE);
    if (d >= n.N) return;
    const uint32_t hi = n.slot_hi[d];
    uint32_t out = n.in_begin[d];
    for (uint32_t k0 = 0; k0 < hi; k0 += BG_WAVE) {
        const uint32_t k = k0 + lane;
        const uint64_t slot = (uint64_t)d * n.M + k;
        const bool use = k < hi && (n.meta[slot] & BG_META_IN_USE);
        const uint64_t ball = __ballot(use);
        if (use) n.in_map[out + __popcll(ball & ((1ull << lane) - 1ull))] = (uint32_t)slot;
        out += (uint32_t)__popcll(ball);
    }
}

__global__ void k_count_outgoing(DevNet n, const uint32_t *in_map, uint64_t total, uint32_t *out_count)
{
    uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (q >= total) return;
    atomicAdd(&out_count[n.meta[in_map[q]] & BG_META_SRC_MASK], 1u);
}

__global__ void k_place_outgoing(DevNet n, const uint32_t *in_map, uint64_t total, const uint32_t *out_begin,
                                 uint32_t *cursor, uint32_t *out_map)
{
    uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (q >= total) return;
    const uint32_t slot = in_map[q];
    const uint32_t s = n.meta[slot] & BG_META_SRC_MASK;
    out_map[out_begin[s] + atomicAdd(&cursor[s], 1u)] = slot;
}

// Sorts each source's outgoing segment ascending (the reference's map is stable in iSyn order).
// Workgroup per source: bitonic sort in LDS for segments up to 4096, serial insertion sort above.
#define BG_SORT_LDS 4096
__global__ __launch_bounds__(BG_BLOCK) void k_sort_outgoing(const uint32_t *out_begin, const uint32_t *out_count,
                                                            uint32_t *out_map)
{
    __shared__ uint32_t v[BG_SORT_LDS];
    const int s = blockIdx.x;
    const uint32_t b = out_begin[s], c = out_count[s];
    if (c < 2) return;
    if (c > BG_SORT_LDS) {
        if (threadIdx.x == 0) {
            for (uint32_t i = 1; i < c; i++) {
                uint32_t x = out_map[b + i];
                uint32_t j = i;
                while (j > 0 && out_map[b + j - 1] > x) { out_map[b + j] = out_map[b + j - 1]; j--; }
                out_map[b + j] = x;
            }
        }
        return;
    }
    uint32_t P = 1;
    while (P < c) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) v[i] = i < c ? out_map[b + i] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                uint32_t l = i ^ j;
                if (l > i) {
                    bool up = (i & k) == 0;
                    uint32_t a = v[i], bb = v[l];
                    if ((a > bb) == up) { v[i] = bb; v[l] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) out_map[b + i] = v[i];
}

// k_flush_block over every slot of every network of a batch (sbase: 64-bit slot prefix).
__global__ __launch_bounds__(BG_BLOCK) void k_flush_batch(const BatchNet *nets, const uint64_t *sbase, int nn,
                                                          const uint64_t *tprev, const int *sel)
{
    const uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (w >= sbase[nn]) return;
    int lo = 0, hi = nn;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (sbase[mid] <= w) lo = mid;
        else hi = mid;
    }
    const DevNet &n = nets[lo].n;
    const uint64_t slot = w - sbase[lo];
    const uint32_t meta = n.meta[slot];
    if (!(meta & BG_META_IN_USE)) return;
    uint32_t fm = n.mask[sel[lo]][meta & BG_META_SRC_MASK];
    if (!fm) return;
    const uint32_t delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
    const uint32_t Rm = (uint32_t)n.R - 1;
    while (fm) {
        const int b0 = __ffs(fm) - 1;
        fm &= fm - 1;
        const uint32_t bit = (uint32_t)((tprev[lo] + (uint64_t)b0 + delay) & Rm);
        uint32_t *word = &n.ring[(uint64_t)(bit >> 5) * n.S + slot];
        const uint32_t m = 1u << (bit & 31);
        if (*word & m) dev_error(n, 1u);
        *word |= m;
    }
}
