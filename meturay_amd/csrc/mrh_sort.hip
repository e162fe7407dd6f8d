// mrh_sort.hip — onesweep LSD radix sort (8-bit digits) and the HitKey
// partition built on it.
//
// One histogram kernel reads the keys once and counts every digit pass;
// a one-block scan turns the counts into global digit bases; then each pass
// is a single kernel: a tile of 4096 keys is ranked in LDS with wave64
// ballots (one ballot per digit bit gives each lane the set of lanes holding
// the same digit), the tile's digit counts are published and the tile looks
// back over its predecessors' published counts (decoupled look-back) to find
// its global offsets, and the keys are staged in LDS in sorted order so the
// global writes of each digit run are contiguous.  Descending order sorts the
// complemented digits, which keeps the sort stable.
#include "mrh_internal.hpp"

namespace mrh {
namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;   // 4096
constexpr int MAX_PASSES = 8;
constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_PREFIX = 2u << 30;
constexpr uint32_t COUNT_MASK = (1u << 30) - 1u;

// meta layout (uint32 words)
constexpr size_t META_HIST = 0;                          // [MAX_PASSES][256]
constexpr size_t META_PART = META_HIST + MAX_PASSES * 256; // [256]
constexpr size_t META_CTR = META_PART + 256;              // [MAX_PASSES]
constexpr size_t META_STATUS = META_CTR + 64;             // [passes][tiles][256]

template <class KeyT>
__device__ __forceinline__ uint32_t digit_of(KeyT k, uint32_t shift, uint32_t mask, uint32_t flip)
{
    return ((uint32_t)(k >> shift) & mask) ^ flip;
}

struct PassDesc {
    uint32_t shift[MAX_PASSES];
    uint32_t mask[MAX_PASSES];
    uint32_t flip[MAX_PASSES];
    uint32_t passes;
    uint32_t part_shift, part_mask, part_flip, part_on;
};

template <class KeyT>
__global__ __launch_bounds__(256) void k_sort_hist(const KeyT* __restrict__ keys, uint32_t n_max,
                                                   const uint32_t* __restrict__ n_dev, PassDesc pd,
                                                   uint32_t* __restrict__ meta)
{
    __shared__ uint32_t s_hist[MAX_PASSES][256];
    __shared__ uint32_t s_part[256];
    for (uint32_t i = threadIdx.x; i < MAX_PASSES * 256; i += blockDim.x) (&s_hist[0][0])[i] = 0;
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_part[i] = 0;
    __syncthreads();
    const uint32_t n = n_dev ? *n_dev : n_max;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        KeyT k = keys[i];
        for (uint32_t p = 0; p < pd.passes; p++)
            atomicAdd(&s_hist[p][digit_of(k, pd.shift[p], pd.mask[p], pd.flip[p])], 1u);
        if (pd.part_on) atomicAdd(&s_part[digit_of(k, pd.part_shift, pd.part_mask, pd.part_flip)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < pd.passes * 256; i += blockDim.x) {
        uint32_t c = (&s_hist[0][0])[i];
        if (c) atomicAdd(&meta[META_HIST + i], c);
    }
    if (pd.part_on)
        for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x)
            if (s_part[i]) atomicAdd(&meta[META_PART + i], s_part[i]);
}

// 256-wide exclusive scan in one block (4 waves, shuffle scan + wave totals)
__device__ __forceinline__ uint32_t block_exclusive_scan256(uint32_t x, uint32_t* s_tmp)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    uint32_t wp = 0;
    for (uint32_t w = 0; w < wave; w++) wp += s_tmp[w];
    __syncthreads();
    return incl + wp - x;
}

__global__ __launch_bounds__(256) void k_sort_scan(uint32_t* __restrict__ meta, uint32_t passes,
                                                   uint32_t* __restrict__ part_offsets,
                                                   uint32_t part_digits, uint32_t n_max,
                                                   const uint32_t* __restrict__ n_dev)
{
    __shared__ uint32_t s_tmp[8];
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t* h = meta + META_HIST + p * 256;
        uint32_t x = h[threadIdx.x];
        uint32_t e = block_exclusive_scan256(x, s_tmp);
        h[threadIdx.x] = e;
    }
    if (part_offsets) {
        uint32_t x = meta[META_PART + threadIdx.x];
        uint32_t e = block_exclusive_scan256(x, s_tmp);
        if (threadIdx.x < part_digits) part_offsets[threadIdx.x] = e;
        if (threadIdx.x == 0) part_offsets[part_digits] = n_dev ? *n_dev : n_max;
    }
}

template <class KeyT>
__global__ __launch_bounds__(SORT_THREADS) void k_onesweep(
    const KeyT* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, KeyT* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n_max, const uint32_t* __restrict__ n_dev, uint32_t shift,
    uint32_t nbits, uint32_t flip, const uint32_t* __restrict__ bases, uint32_t* __restrict__ status,
    uint32_t* __restrict__ tile_ctr)
{
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_whist[SORT_WAVES][256];
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_tmp[8];
    __shared__ KeyT s_keys[SORT_TILE];
    __shared__ uint32_t s_vals[SORT_TILE];

    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t mask = (1u << nbits) - 1u;
    const uint32_t n = n_dev ? *n_dev : n_max;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (uint32_t i = tid; i < SORT_WAVES * 256; i += SORT_THREADS) (&s_whist[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t tile_start = tile * SORT_TILE;
    if (tile_start >= n) return;   // same for the whole block
    const uint32_t tile_n = min((uint32_t)SORT_TILE, n - tile_start);

    KeyT k[SORT_ITEMS];
    uint32_t v[SORT_ITEMS], rk[SORT_ITEMS], dg[SORT_ITEMS];
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    const uint32_t wbase = wave * 64 * SORT_ITEMS;
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
        const uint32_t local = wbase + j * 64 + lane;
        const bool valid = local < tile_n;
        const uint32_t g = tile_start + local;
        KeyT key = valid ? keys_in[g] : (KeyT)0;
        uint32_t val = valid ? (vals_in ? vals_in[g] : g) : 0u;
        uint32_t d = digit_of(key, shift, mask, flip);
        uint64_t peers = __ballot(valid);
        for (uint32_t b = 0; b < nbits; b++) {
            uint32_t bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t pre = (uint32_t)__popcll(peers & lanes_below);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t base = valid ? s_whist[wave][d] : 0u;
        if (valid && pre == 0) s_whist[wave][d] = base + cnt;
        k[j] = key;
        v[j] = val;
        rk[j] = base + pre;
        dg[j] = valid ? d : 0xFFFFFFFFu;
    }
    __syncthreads();

    // exclusive offsets of each wave inside a digit, tile total per digit
    const uint32_t d = tid;
    uint32_t count = 0;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; w++) {
        uint32_t t = s_whist[w][d];
        s_whist[w][d] = count;
        count += t;
    }
    s_start[d] = block_exclusive_scan256(count, s_tmp);

    // decoupled look-back: publish this tile's count, add predecessors'
    uint32_t* st = status + (size_t)tile * 256;
    uint32_t excl = 0;
    if (tile == 0) {
        __hip_atomic_store(&st[d], FLAG_PREFIX | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(&st[d], FLAG_AGG | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int t = (int)tile - 1;
        for (;;) {
            uint32_t s = __hip_atomic_load(&status[(size_t)t * 256 + d], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            if ((s >> 30) == 0u) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += s & COUNT_MASK;
            if ((s & FLAG_PREFIX) || t == 0) break;
            t--;
        }
        __hip_atomic_store(&st[d], FLAG_PREFIX | (excl + count), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gbase[d] = bases[d] + excl;
    __syncthreads();

#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
        if (dg[j] != 0xFFFFFFFFu) {
            uint32_t pos = s_start[dg[j]] + s_whist[wave][dg[j]] + rk[j];
            s_keys[pos] = k[j];
            s_vals[pos] = v[j];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < tile_n; i += SORT_THREADS) {
        KeyT key = s_keys[i];
        uint32_t dd = digit_of(key, shift, mask, flip);
        uint32_t o = s_gbase[dd] + (i - s_start[dd]);
        if (keys_out) keys_out[o] = key;
        vals_out[o] = s_vals[i];
    }
}

template <class KeyT>
int32_t radix_sort_t(mrh_context* ctx, const SortArgs& a, const SortSpace* space)
{
    const uint32_t nbits_total = a.end_bit - a.begin_bit;
    const uint32_t passes = (nbits_total + 7) / 8;
    const uint32_t tiles = div_up(a.n_max, SORT_TILE);
    PassDesc pd = {};
    pd.passes = passes;
    for (uint32_t p = 0; p < passes; p++) {
        uint32_t sh = a.begin_bit + 8 * p;
        uint32_t nb = min(8u, a.end_bit - sh);
        pd.shift[p] = sh;
        pd.mask[p] = (1u << nb) - 1u;
        pd.flip[p] = a.descending ? pd.mask[p] : 0u;
    }
    uint32_t part_digits = 0;
    if (a.part_offsets) {
        uint32_t w = a.end_bit - a.part_begin;
        part_digits = 1u << w;
        pd.part_on = 1;
        pd.part_shift = a.part_begin;
        pd.part_mask = part_digits - 1u;
        pd.part_flip = a.descending ? pd.part_mask : 0u;
    }

    const size_t meta_words = META_STATUS + (size_t)passes * tiles * 256;
    const bool plain = space != nullptr;   // a fixed workspace: no allocation, no events
    uint32_t* meta = nullptr;
    void *ka = nullptr, *kb = nullptr, *va = nullptr, *vb = nullptr;
    if (space) {
        if (meta_words > space->meta_words || a.n_max > space->max_n ||
            (passes > 1 && (!space->keys_a || !space->vals_a)) ||
            (passes > 2 && (!space->keys_b || !space->vals_b)))
            return fail(ctx, MRH_ERR_INVALID, "radix_sort: %u keys over %u passes exceed the workspace",
                        a.n_max, passes);
        meta = space->meta;
        ka = space->keys_a;
        kb = space->keys_b;
        va = space->vals_a;
        vb = space->vals_b;
    } else {
        void* meta_v = nullptr;
        MRH_TRY(scratch(ctx, SLOT_SORT_META, meta_words * 4, &meta_v));
        meta = (uint32_t*)meta_v;
        // ping-pong buffers for intermediate passes
        if (passes > 1) {
            MRH_TRY(scratch(ctx, SLOT_SORT_KEYS_A, (size_t)a.n_max * sizeof(KeyT), &ka));
            MRH_TRY(scratch(ctx, SLOT_SORT_VALS_A, (size_t)a.n_max * 4, &va));
            if (passes > 2) {
                MRH_TRY(scratch(ctx, SLOT_SORT_KEYS_B, (size_t)a.n_max * sizeof(KeyT), &kb));
                MRH_TRY(scratch(ctx, SLOT_SORT_VALS_B, (size_t)a.n_max * 4, &vb));
            }
        }
    }
    // histograms, counters and look-back words start at zero on every call
    MRH_TRY(fill_u32(ctx, meta, 0u, meta_words, plain));

    // Each block flushes passes x 256 counters with global atomics, so the
    // grid is sized for >= 4096 keys per block and at most 4 blocks per CU:
    // one block per 256 keys put millions of atomics on 1024 words.
    uint32_t hist_blocks = div_up(a.n_max, 4096u);
    hist_blocks = std::min<uint32_t>(hist_blocks, ctx->num_cus * 4u);
    hist_blocks = std::max<uint32_t>(hist_blocks, 1u);
    auto run = [&](const char* name, dim3 grid, auto kernel, auto... args) -> int32_t {
        return plain ? launch_plain(ctx, name, grid, dim3(256), 0, kernel, args...)
                     : launch(ctx, name, grid, dim3(256), 0, kernel, args...);
    };
    static_assert(SORT_THREADS == 256, "every sort kernel runs 256-thread blocks");
    MRH_TRY(run("sort_hist", dim3(hist_blocks), k_sort_hist<KeyT>, (const KeyT*)a.keys_in, a.n_max,
                a.n_dev, pd, meta));
    MRH_TRY(run("sort_scan", dim3(1), k_sort_scan, meta, passes, a.part_offsets, part_digits, a.n_max,
                a.n_dev));

    const KeyT* kin = (const KeyT*)a.keys_in;
    const uint32_t* vin = a.vals_in;
    for (uint32_t p = 0; p < passes; p++) {
        const bool last = (p + 1 == passes);
        KeyT* kout = last ? (KeyT*)a.keys_out : (KeyT*)((p & 1) ? kb : ka);
        uint32_t* vout = last ? a.vals_out : (uint32_t*)((p & 1) ? vb : va);
        MRH_TRY(run("sort_onesweep", dim3(tiles), k_onesweep<KeyT>, kin, vin, kout, vout, a.n_max,
                    a.n_dev, pd.shift[p], (uint32_t)__builtin_popcount(pd.mask[p]), pd.flip[p],
                    (const uint32_t*)(meta + META_HIST + p * 256),
                    meta + META_STATUS + (size_t)p * tiles * 256, meta + META_CTR + p));
        kin = kout;
        vin = vout;
    }
    return MRH_OK;
}

// Fixed workspace of mrh_partition_by_key_noalloc: up to 2^22 keys (32-bit
// keys, so at most 4 passes): the meta words, then keys A, values A, keys B,
// values B of NOALLOC_MAX_KEYS words each.
constexpr uint32_t NOALLOC_MAX_KEYS = MRH_PARTITION_NOALLOC_MAX_KEYS;
constexpr size_t NOALLOC_META_WORDS =
    (META_STATUS + (size_t)4 * (NOALLOC_MAX_KEYS / SORT_TILE) * 256 + 63) / 64 * 64;
constexpr size_t NOALLOC_WORDS = NOALLOC_META_WORDS + (size_t)4 * NOALLOC_MAX_KEYS;
static_assert(NOALLOC_MAX_KEYS % SORT_TILE == 0, "whole tiles");

}  // namespace
}  // namespace mrh

// One copy per device, held until the process ends.  Host code never uses
// this name as an address: noalloc_setup looks up the device address of
// each device's copy, and the kernels receive that.
static __device__ __align__(256) uint32_t scratch[mrh::NOALLOC_WORDS];

namespace mrh {

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

int32_t noalloc_setup(mrh_context* ctx)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) {
        ctx->noalloc_err = "device index beyond the fixed workspace table";
        return MRH_ERR_UNSUPPORTED;
    }
    void*& p = g_noalloc_ptr[ctx->device];
    if (!p) {
        void* addr = nullptr;
        hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || !addr) {
            (void)hipGetLastError();
            ctx->noalloc_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
            return MRH_ERR_HIP;
        }
        p = addr;
    }
    ctx->noalloc_scratch = p;
    return MRH_OK;
}

int32_t radix_sort(mrh_context* ctx, const SortArgs& a, const SortSpace* space)
{
    if (a.n_max == 0) return MRH_OK;
    if (a.n_max >= (1u << 30))
        return fail(ctx, MRH_ERR_INVALID, "radix_sort: n = %u exceeds 2^30 - 1", a.n_max);
    if (a.end_bit <= a.begin_bit || a.end_bit > a.key_bytes * 8u)
        return fail(ctx, MRH_ERR_INVALID, "radix_sort: bad bit range [%u, %u)", a.begin_bit,
                    a.end_bit);
    if ((a.end_bit - a.begin_bit + 7) / 8 > (uint32_t)MAX_PASSES)
        return fail(ctx, MRH_ERR_INVALID, "radix_sort: too many passes");
    if (!a.vals_out) return fail(ctx, MRH_ERR_INVALID, "radix_sort: vals_out is required");
    if (a.part_offsets && (a.end_bit - a.part_begin > 8 || a.part_begin < a.begin_bit))
        return fail(ctx, MRH_ERR_INVALID, "radix_sort: partition bits must be <= 8 wide and inside the sorted range");
    if (a.key_bytes == 4) return radix_sort_t<uint32_t>(ctx, a, space);
    if (a.key_bytes == 8) return radix_sort_t<uint64_t>(ctx, a, space);
    return fail(ctx, MRH_ERR_INVALID, "radix_sort: key_bytes must be 4 or 8");
}

}  // namespace mrh

using namespace mrh;

extern "C" {

int32_t mrh_sort_pairs(mrh_context* ctx, const void* d_keys_in, const uint32_t* d_vals_in,
                       void* d_keys_out, uint32_t* d_vals_out, uint32_t n, uint32_t key_bytes,
                       uint32_t begin_bit, uint32_t end_bit, int32_t descending)
{
    if (!ctx) return MRH_ERR_INVALID;
    if (n && (!d_keys_in || !d_vals_out))
        return fail(ctx, MRH_ERR_INVALID, "mrh_sort_pairs: null input/output");
    SortArgs a;
    a.keys_in = d_keys_in;
    a.vals_in = d_vals_in;
    a.keys_out = d_keys_out;
    a.vals_out = d_vals_out;
    a.n_max = n;
    a.key_bytes = key_bytes;
    a.begin_bit = begin_bit;
    a.end_bit = end_bit;
    a.descending = descending != 0;
    return radix_sort(ctx, a);
}

int32_t mrh_partition_by_key(mrh_context* ctx, const uint32_t* d_keys, uint32_t n,
                             uint32_t sort_begin_bit, uint32_t part_begin_bit, uint32_t end_bit,
                             uint32_t* d_order, uint32_t* d_offsets)
{
    if (!ctx) return MRH_ERR_INVALID;
    if (!d_order || !d_offsets || (n && !d_keys))
        return fail(ctx, MRH_ERR_INVALID, "mrh_partition_by_key: null argument");
    if (end_bit > 32 || part_begin_bit < sort_begin_bit || end_bit <= part_begin_bit ||
        end_bit - part_begin_bit > 8)
        return fail(ctx, MRH_ERR_INVALID, "mrh_partition_by_key: bad bit range");
    if (n == 0) {
        uint32_t digits = 1u << (end_bit - part_begin_bit);
        return fill_u32(ctx, d_offsets, 0u, digits + 1u);
    }
    SortArgs a;
    a.keys_in = d_keys;
    a.keys_out = nullptr;
    a.vals_out = d_order;
    a.n_max = n;
    a.key_bytes = 4;
    a.begin_bit = sort_begin_bit;
    a.end_bit = end_bit;
    a.descending = true;
    a.part_offsets = d_offsets;
    a.part_begin = part_begin_bit;
    // intermediate passes need key storage even when the caller wants none
    if ((end_bit - sort_begin_bit + 7) / 8 > 1) {
        void* kout = nullptr;
        MRH_TRY(mrh::scratch(ctx, SLOT_PART_KEYS, (size_t)n * 4, &kout));
        a.keys_out = kout;
    }
    return radix_sort(ctx, a);
}

int32_t mrh_partition_by_key_noalloc(mrh_context* ctx, const uint32_t* d_keys, uint32_t n,
                                     uint32_t sort_begin_bit, uint32_t part_begin_bit,
                                     uint32_t end_bit, uint32_t* d_order, uint32_t* d_offsets)
{
    if (!ctx) return MRH_ERR_INVALID;
    if (!d_order || !d_offsets || (n && !d_keys))
        return fail(ctx, MRH_ERR_INVALID, "mrh_partition_by_key_noalloc: null argument");
    if (end_bit > 32 || part_begin_bit < sort_begin_bit || end_bit <= part_begin_bit ||
        end_bit - part_begin_bit > 8)
        return fail(ctx, MRH_ERR_INVALID, "mrh_partition_by_key_noalloc: bad bit range");
    if (n > NOALLOC_MAX_KEYS)
        return fail(ctx, MRH_ERR_INVALID,
                    "mrh_partition_by_key_noalloc: n = %u exceeds the fixed workspace of %u keys "
                    "(mrh_partition_by_key has no limit)", n, NOALLOC_MAX_KEYS);
    if (!ctx->noalloc_scratch)
        return fail(ctx, MRH_ERR_HIP, "mrh_partition_by_key_noalloc: no workspace on device %d: %s",
                    ctx->device, ctx->noalloc_err.c_str());
    if (n == 0) {
        uint32_t digits = 1u << (end_bit - part_begin_bit);
        return fill_u32(ctx, d_offsets, 0u, digits + 1u, true);
    }
    uint32_t* base = (uint32_t*)ctx->noalloc_scratch;
    SortSpace ws;
    ws.meta = base;
    ws.meta_words = NOALLOC_META_WORDS;
    ws.keys_a = base + NOALLOC_META_WORDS;
    ws.vals_a = base + NOALLOC_META_WORDS + (size_t)NOALLOC_MAX_KEYS;
    ws.keys_b = base + NOALLOC_META_WORDS + (size_t)2 * NOALLOC_MAX_KEYS;
    ws.vals_b = base + NOALLOC_META_WORDS + (size_t)3 * NOALLOC_MAX_KEYS;
    ws.max_n = NOALLOC_MAX_KEYS;
    SortArgs a;
    a.keys_in = d_keys;
    a.keys_out = nullptr;     // the last pass writes only the order
    a.vals_out = d_order;
    a.n_max = n;
    a.key_bytes = 4;
    a.begin_bit = sort_begin_bit;
    a.end_bit = end_bit;
    a.descending = true;
    a.part_offsets = d_offsets;
    a.part_begin = part_begin_bit;
    return radix_sort(ctx, a, &ws);
}

}  // extern "C"
