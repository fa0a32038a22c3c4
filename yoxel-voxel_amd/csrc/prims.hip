// Device-wide scans, a stable LSD radix sort and select-distinct (gfx950).
//
// Scans are reduce-then-scan in three launches: per-block sums of 4096 items
// (256 threads, 16 coalesced rounds), one 1024-thread block that scans the
// block sums (a strip per thread), and a pass that rescans each block's items
// with its carry. Items come through a load functor and results leave
// through a store functor, so flag scans, widening scans and the
// select-distinct scatter need no intermediate arrays.
//
// The radix sort takes 8-bit digits from the least significant end: per pass
// a digit histogram per block (a block takes a run of tiles of 4096 keys; the
// histogram is stored digit-major, so one exclusive scan of it gives every
// (digit, block) pair its first output position) and a scatter that ranks
// equal digits in key order with wave ballots, so every pass is stable.
#include "prims.h"

#include <algorithm>

#include "internal.h"

namespace yx {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanRounds = (int)(kPrimScanChunk / kScanThreads);
constexpr int kSortThreads = 256;
constexpr int kSortRounds = (int)(kPrimSortTile / kSortThreads);

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- scan ----

template <class Src, class Acc>
__global__ __launch_bounds__(kScanThreads) void prim_scan_reduce(Src src, uint64_t n, Acc* partial) {
    __shared__ Acc shw[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kPrimScanChunk + threadIdx.x;
    Acc s = 0;
#pragma unroll
    for (int k = 0; k < kScanRounds; ++k) {
        const uint64_t i = base + (uint64_t)k * kScanThreads;
        s += i < n ? (Acc)src(i) : (Acc)0;
    }
    Acc tot;
    (void)block_excl_scan<Acc, kScanThreads>(s, shw, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// In-place exclusive scan of the np block sums; *total gets their sum.
template <class Acc, class TotalT>
__global__ __launch_bounds__(1024) void prim_scan_partials(Acc* partial, uint32_t np, TotalT* total) {
    __shared__ Acc shw[1024 / 64];
    const uint32_t per = (np + 1023u) / 1024u;
    const uint32_t b = std::min(threadIdx.x * per, np), e = std::min(b + per, np);
    Acc s = 0;
    for (uint32_t i = b; i < e; ++i) s += partial[i];
    Acc tot;
    Acc run = block_excl_scan<Acc, 1024>(s, shw, tot);
    for (uint32_t i = b; i < e; ++i) {
        const Acc v = partial[i];
        partial[i] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total) *total = (TotalT)tot;
}

// dst(i, exclusive prefix, item) for every item of the block's chunk. A
// thread reads its item before the block scan and stores after it, so the
// store may overwrite the loaded array in place.
template <class Src, class Dst, class Acc>
__global__ __launch_bounds__(kScanThreads) void prim_scan_apply(Src src, Dst dst, uint64_t n, const Acc* partial) {
    __shared__ Acc shw[kScanThreads / 64];
    Acc carry = partial[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kPrimScanChunk + threadIdx.x;
    for (int k = 0; k < kScanRounds; ++k) {
        const uint64_t i = base + (uint64_t)k * kScanThreads;
        const Acc v = i < n ? (Acc)src(i) : (Acc)0;
        Acc tot;
        const Acc ex = block_excl_scan<Acc, kScanThreads>(v, shw, tot);
        if (i < n) dst(i, carry + ex, v);
        carry += tot;
    }
}

// Scans of at most kScanSingleMax items run as one 1024-thread block (a strip
// of at most 8 items per thread: sums, one block scan, rescans with the carry)
// instead of three launches: small scans are bound by launch and barrier
// latency, not bandwidth.
constexpr uint32_t kScanSingleMax = 8u * 1024u;

template <class Src, class Dst, class Acc, class TotalT>
__global__ __launch_bounds__(1024) void prim_scan_single(Src src, Dst dst, uint32_t n, bool apply, TotalT* total) {
    __shared__ Acc shw[1024 / 64];
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = std::min(threadIdx.x * per, n), e = std::min(b + per, n);
    Acc s = 0;
    for (uint32_t i = b; i < e; ++i) s += (Acc)src(i);
    Acc tot;
    Acc run = block_excl_scan<Acc, 1024>(s, shw, tot);
    if (apply) {
        for (uint32_t i = b; i < e; ++i) {
            const Acc v = (Acc)src(i);
            dst(i, run, v);
            run += v;
        }
    }
    if (threadIdx.x == 0 && total) *total = (TotalT)tot;
}

template <class T>
struct Load {
    const T* p;
    __device__ T operator()(uint64_t i) const { return p[i]; }
};

template <class T>
struct StorePrefix {
    T* p;
    __device__ void operator()(uint64_t i, T prefix, T) const { p[i] = prefix; }
};

struct NoStore {
    template <class T>
    __device__ void operator()(uint64_t, T, T) const {}
};

// 1 where the shifted value differs from its predecessor's (and at 0).
struct FirstOfRun {
    const uint64_t* p;
    int shift;
    __device__ uint32_t operator()(uint64_t i) const {
        return (i == 0 || (p[i] >> shift) != (p[i - 1] >> shift)) ? 1u : 0u;
    }
};

struct StoreFirst {
    const uint64_t* p;
    uint64_t* out;
    int shift;
    __device__ void operator()(uint64_t i, uint32_t prefix, uint32_t first) const {
        if (first) out[prefix] = p[i] >> shift;
    }
};

template <class Acc, class Src, class Dst, class TotalT>
int scan_run(Src src, Dst dst, bool apply, uint64_t n, TotalT* total, void* ws, size_t ws_bytes, hipStream_t st) {
    if (n >= (1ull << 32)) return fail(YX_ERR_RANGE, "scan of %llu items (at most 2^32 - 1)", (unsigned long long)n);
    if (!ws || ws_bytes < scan_workspace_bytes(n)) return fail(YX_ERR_INVALID_ARG, "scan workspace too small");
    if (n <= kScanSingleMax) {
        hipLaunchKernelGGL((prim_scan_single<Src, Dst, Acc, TotalT>), dim3(1), dim3(1024), 0, st, src, dst,
                           (uint32_t)n, apply, total);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(YX_ERR_HIP, "scan launch: %s", hipGetErrorString(e));
        return YX_OK;
    }
    const uint32_t np = (uint32_t)((n + kPrimScanChunk - 1) / kPrimScanChunk);
    Acc* partial = (Acc*)ws;
    if (np) hipLaunchKernelGGL((prim_scan_reduce<Src, Acc>), dim3(np), dim3(kScanThreads), 0, st, src, n, partial);
    hipLaunchKernelGGL((prim_scan_partials<Acc, TotalT>), dim3(1), dim3(1024), 0, st, partial, np, total);
    if (apply && np)
        hipLaunchKernelGGL((prim_scan_apply<Src, Dst, Acc>), dim3(np), dim3(kScanThreads), 0, st, src, dst, n,
                           (const Acc*)partial);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(YX_ERR_HIP, "scan launch: %s", hipGetErrorString(e));
    return YX_OK;
}

// ---- radix sort ----

// The n keys are cut into tiles of 4096; block b of nb takes the consecutive
// tiles [b * per, (b + 1) * per), per = ceil(tiles / nb), computed from the
// count the kernels read (so a count known only on the device still spreads
// evenly over the blocks). nb is bounded (see resident_sort_blocks): the
// digit-major histogram has 256 * nb entries, and each block writes and reads
// its 256 entries scattered over it.
__device__ __forceinline__ void sort_block_tiles(uint32_t n, uint32_t nb, uint32_t& first, uint32_t& last) {
    const uint32_t tiles = (n + kPrimSortTile - 1) / kPrimSortTile;
    const uint32_t per = (tiles + nb - 1) / nb;
    first = std::min(blockIdx.x * per, tiles);
    last = std::min(first + per, tiles);
}

// Digit histogram of block b's keys, stored digit-major (hist[d * nb + b]);
// blocks without keys store zeros. Each run of equal digits among a wave's
// lanes is counted with one LDS add by its first lane (keys that arrive
// grouped, such as records emitted tile by tile, share digits with their
// neighbours; per-lane adds to one counter would serialise).
// VEC (keys 16-byte aligned): each thread loads 16 consecutive bytes of keys
// per round and adds each run of equal digits among them with one LDS add.
template <class K, bool VEC>
__global__ __launch_bounds__(kSortThreads) void prim_rs_upsweep(const K* keys, uint32_t n_host, const uint32_t* n_dev,
                                                                int shift, int bits, uint32_t* hist, uint32_t nb) {
    __shared__ uint32_t h[kPrimRadix];
    for (int d = (int)threadIdx.x; d < kPrimRadix; d += kSortThreads) h[d] = 0u;
    __syncthreads();
    const uint32_t n = n_dev ? std::min(*n_dev, n_host) : n_host;
    const uint32_t mask = (1u << bits) - 1u;
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t above = lane == 63 ? 0ull : (~0ull << (lane + 1));
    uint32_t t0, t1;
    sort_block_tiles(n, nb, t0, t1);
    if constexpr (VEC) {
        constexpr int E = 16 / (int)sizeof(K);                             // keys per load
        constexpr int R = (int)(kPrimSortTile / (kSortThreads * E));       // loads per tile
        struct alignas(16) Vec {
            K k[E];
        };
        constexpr uint32_t kNone = 0xFFFFFFFFu;  // no key (beyond n)
        for (uint32_t tile = t0; tile < t1; ++tile) {
            uint32_t dr[R][E];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t i0 = tile * kPrimSortTile + ((uint32_t)r * kSortThreads + threadIdx.x) * (uint32_t)E;
                if (i0 + (uint32_t)E <= n) {
                    const Vec v = *reinterpret_cast<const Vec*>(keys + i0);
#pragma unroll
                    for (int j = 0; j < E; ++j) dr[r][j] = (uint32_t)(v.k[j] >> shift) & mask;
                } else {
#pragma unroll
                    for (int j = 0; j < E; ++j)
                        dr[r][j] = i0 + (uint32_t)j < n ? (uint32_t)(keys[i0 + j] >> shift) & mask : kNone;
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                uint32_t cur = dr[r][0], cnt = 1u;
#pragma unroll
                for (int j = 1; j < E; ++j) {
                    if (dr[r][j] == cur) {
                        ++cnt;
                    } else {
                        if (cur != kNone) atomicAdd(&h[cur], cnt);
                        cur = dr[r][j];
                        cnt = 1u;
                    }
                }
                if (cur != kNone) atomicAdd(&h[cur], cnt);
            }
        }
    } else
    for (uint32_t tile = t0; tile < t1; ++tile) {
        const uint32_t base = tile * kPrimSortTile + threadIdx.x;
        // every round's digit first: the block's 16 loads per thread are in
        // flight together instead of one round's at a time
        uint32_t dr[kSortRounds];
#pragma unroll
        for (int k = 0; k < kSortRounds; ++k) {
            const uint32_t i = base + (uint32_t)k * kSortThreads;
            dr[k] = i < n ? (uint32_t)(keys[i] >> shift) & mask : 0u;
        }
#pragma unroll
        for (int k = 0; k < kSortRounds; ++k) {
            const uint32_t i = base + (uint32_t)k * kSortThreads;
            const bool in = i < n;  // the lanes in range are a prefix of the wave
            const uint32_t d = dr[k];
            const uint32_t prev = __shfl_up(d, 1u);
            const bool head = in && (lane == 0 || prev != d);
            const uint64_t heads = __ballot(head);
            const int n_in = __popcll(__ballot(in));
            if (head) {
                const uint64_t nxt = heads & above;
                const int end = nxt ? __ffsll((unsigned long long)nxt) - 1 : n_in;
                atomicAdd(&h[d], (uint32_t)(end - lane));
            }
        }
    }
    __syncthreads();
    for (int d = (int)threadIdx.x; d < kPrimRadix; d += kSortThreads) hist[(size_t)d * nb + blockIdx.x] = h[d];
}

// Stable scatter of block b's keys (and ids) by digit; `offs` is the scanned
// histogram. Each wave of the block takes one contiguous quarter of a tile
// (16 coalesced rounds of 64 keys) and ranks its keys among its own keys of
// the same digit: earlier rounds through a per-wave count in LDS, the same
// round by ballot rank. One barrier later, a thread per digit turns the four
// waves' counts into their bases (the block's keys of that digit from earlier
// tiles, then the earlier waves' of this tile), and after a second barrier
// every lane writes its 16 keys. Two barriers per tile of 4096 keys.
// IDS with ids_in null: the ids are the positions.
template <class K, bool IDS>
__global__ __launch_bounds__(kSortThreads) void prim_rs_downsweep(const K* keys_in, const uint32_t* ids_in,
                                                                  K* keys_out, uint32_t* ids_out, uint32_t n_host,
                                                                  const uint32_t* n_dev, int shift, int bits,
                                                                  const uint32_t* offs, uint32_t nb) {
    constexpr int kWaves = kSortThreads / 64;
    static_assert(kSortThreads == kPrimRadix, "one thread per digit in the base step");
    __shared__ uint32_t run[kPrimRadix];             // the block's keys per digit so far (with the offset)
    __shared__ uint32_t wcnt[kWaves][kPrimRadix];    // this tile's keys per wave and digit
    __shared__ uint32_t wbase[kWaves][kPrimRadix];   // where each wave's keys of a digit start
    const uint32_t n = n_dev ? std::min(*n_dev, n_host) : n_host;
    uint32_t t0, t1;
    sort_block_tiles(n, nb, t0, t1);
    if (t0 >= t1) return;  // whole block
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const uint32_t mask = (1u << bits) - 1u;
    {
        const int d = (int)threadIdx.x;
        run[d] = offs[(size_t)d * nb + blockIdx.x];
        for (int q = 0; q < kWaves; ++q) wcnt[q][d] = 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t tile = t0; tile < t1; ++tile) {
        const uint32_t base = tile * kPrimSortTile + (uint32_t)w * (kSortRounds * 64) + (uint32_t)lane;
        // every round's key and id first, so the loads overlap
        K kr[kSortRounds];
        uint32_t ir[IDS ? kSortRounds : 1];
#pragma unroll
        for (int k = 0; k < kSortRounds; ++k) {
            const uint32_t i = base + (uint32_t)k * 64u;
            kr[k] = i < n ? keys_in[i] : (K)0;
            if (IDS) ir[k] = i < n ? (ids_in ? ids_in[i] : i) : 0u;
        }
        uint32_t rk[kSortRounds];
#pragma unroll
        for (int k = 0; k < kSortRounds; ++k) {
            const uint32_t i = base + (uint32_t)k * 64u;
            const bool in = i < n;
            const uint32_t d = in ? (uint32_t)(kr[k] >> shift) & mask : 0u;
            const uint64_t peers = digit_peers(in, d, bits);
            const uint32_t before = wcnt[w][d];
            rk[k] = before + (uint32_t)__popcll(peers & below);
            // the lowest lane of each digit group advances the wave's count;
            // the next round's read sees it (a wave's LDS accesses are in order)
            if (in && (peers & below) == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
        }
        __syncthreads();
        {
            const int d = (int)threadIdx.x;
            uint32_t s = run[d];
#pragma unroll
            for (int q = 0; q < kWaves; ++q) {
                wbase[q][d] = s;
                s += wcnt[q][d];
                wcnt[q][d] = 0u;  // for the next tile (its counts start after the next barrier)
            }
            run[d] = s;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortRounds; ++k) {
            const uint32_t i = base + (uint32_t)k * 64u;
            if (i < n) {
                const uint32_t d = (uint32_t)(kr[k] >> shift) & mask;
                const uint32_t pos = wbase[w][d] + rk[k];
                keys_out[pos] = kr[k];
                if (IDS) ids_out[pos] = ir[k];
            }
        }
    }
}

// Blocks of a sort: one per tile, at most kSortBlocksPerCU per compute unit
// of the current device, and never more than kSortMaxBlocks, which sizes the
// workspace. Fewer blocks mean fewer scattered histogram entries (the
// upsweep's 256 stores and the downsweep's 256 loads per block), more mean
// better balance; on composite_512's 82 M records 16 per CU (about 5 tiles a
// block) measured best of 4, 6, 8, 12, 16, 24 and 32.
constexpr uint32_t kSortBlocksPerCU = 16;
constexpr uint32_t kSortMaxBlocks = 8192;
uint32_t resident_sort_blocks() {
    static int cus_of[64] = {};  // compute units per device, filled on first use (same value on a race)
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
    if (cus_of[dev] <= 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cus_of[dev] = cus;
    }
    return std::min<uint32_t>(kSortMaxBlocks, (uint32_t)cus_of[dev] * kSortBlocksPerCU);
}
uint32_t sort_tiles(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kPrimSortTile - 1) / kPrimSortTile); }

}  // namespace

size_t scan_workspace_bytes(uint64_t n) { return align256(((n + kPrimScanChunk - 1) / kPrimScanChunk + 1) * 8); }

int scan_excl_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* ws, size_t ws_bytes,
                  hipStream_t st) {
    return scan_run<uint32_t>(Load<uint32_t>{in}, StorePrefix<uint32_t>{out}, true, n, total, ws, ws_bytes, st);
}

int scan_excl_u8(const uint8_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* ws, size_t ws_bytes,
                 hipStream_t st) {
    return scan_run<uint32_t>(Load<uint8_t>{in}, StorePrefix<uint32_t>{out}, true, n, total, ws, ws_bytes, st);
}

int scan_excl_u32_u64(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total, void* ws, size_t ws_bytes,
                      hipStream_t st) {
    return scan_run<uint64_t>(Load<uint32_t>{in}, StorePrefix<uint64_t>{out}, true, n, total, ws, ws_bytes, st);
}

int reduce_sum_u32(const uint32_t* in, uint64_t n, uint32_t* total, void* ws, size_t ws_bytes, hipStream_t st) {
    return scan_run<uint32_t>(Load<uint32_t>{in}, NoStore{}, false, n, total, ws, ws_bytes, st);
}

size_t sort_workspace_bytes(uint64_t n_host) {
    const uint64_t hist = (uint64_t)kPrimRadix * std::min<uint64_t>(sort_tiles(n_host), kSortMaxBlocks);
    return align256(hist * 4) + scan_workspace_bytes(hist);
}

template <class K>
int radix_sort(const K* keys_in, const uint32_t* ids_in, K* const kbuf[2], uint32_t* const ibuf[2], uint64_t n_host,
               const uint32_t* n_dev, int b0, int b1, const K** kout, const uint32_t** iout, void* ws,
               size_t ws_bytes, hipStream_t st) {
    if (n_host >= 0x80000000ull) return fail(YX_ERR_RANGE, "sort of %llu keys (at most 2^31 - 1)",
                                             (unsigned long long)n_host);
    if (!ws || ws_bytes < sort_workspace_bytes(n_host)) return fail(YX_ERR_INVALID_ARG, "sort workspace too small");
    if (b0 < 0 || b1 > (int)(8 * sizeof(K))) return fail(YX_ERR_INVALID_ARG, "sort bits [%d, %d)", b0, b1);
    const uint32_t nb = std::min(sort_tiles(n_host), resident_sort_blocks());
    const size_t nhist = (size_t)kPrimRadix * nb;
    uint32_t* hist = (uint32_t*)ws;
    char* scan_ws = (char*)ws + align256(nhist * 4);
    const size_t scan_bytes = ws_bytes - align256(nhist * 4);
    const bool ids = ibuf != nullptr;
    const K* kin = keys_in;
    const uint32_t* iin = ids_in;
    int pass = 0;
    for (int sh = b0; sh < b1; sh += kPrimRadixBits, ++pass) {
        const int bits = std::min(kPrimRadixBits, b1 - sh);
        if (((uintptr_t)kin & 15u) == 0)
            hipLaunchKernelGGL((prim_rs_upsweep<K, true>), dim3(nb), dim3(kSortThreads), 0, st, kin, (uint32_t)n_host,
                               n_dev, sh, bits, hist, nb);
        else
            hipLaunchKernelGGL((prim_rs_upsweep<K, false>), dim3(nb), dim3(kSortThreads), 0, st, kin, (uint32_t)n_host,
                               n_dev, sh, bits, hist, nb);
        const int rc = scan_excl_u32(hist, hist, nhist, nullptr, scan_ws, scan_bytes, st);
        if (rc != YX_OK) return rc;
        if (ids)
            hipLaunchKernelGGL((prim_rs_downsweep<K, true>), dim3(nb), dim3(kSortThreads), 0, st, kin, iin,
                               kbuf[pass & 1], ibuf[pass & 1], (uint32_t)n_host, n_dev, sh, bits,
                               (const uint32_t*)hist, nb);
        else
            hipLaunchKernelGGL((prim_rs_downsweep<K, false>), dim3(nb), dim3(kSortThreads), 0, st, kin,
                               (const uint32_t*)nullptr, kbuf[pass & 1], (uint32_t*)nullptr, (uint32_t)n_host, n_dev,
                               sh, bits, (const uint32_t*)hist, nb);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(YX_ERR_HIP, "radix sort launch: %s", hipGetErrorString(e));
        kin = kbuf[pass & 1];
        iin = ids ? ibuf[pass & 1] : nullptr;
    }
    *kout = kin;
    if (iout) *iout = iin;
    return YX_OK;
}

template int radix_sort<uint32_t>(const uint32_t*, const uint32_t*, uint32_t* const[2], uint32_t* const[2], uint64_t,
                                  const uint32_t*, int, int, const uint32_t**, const uint32_t**, void*, size_t,
                                  hipStream_t);
template int radix_sort<uint64_t>(const uint64_t*, const uint32_t*, uint64_t* const[2], uint32_t* const[2], uint64_t,
                                  const uint32_t*, int, int, const uint64_t**, const uint32_t**, void*, size_t,
                                  hipStream_t);

size_t distinct_workspace_bytes(uint64_t n) { return scan_workspace_bytes(n); }

int distinct_shifted_u64(const uint64_t* in, uint64_t* out, uint64_t n, int shift, uint64_t* count, void* ws,
                         size_t ws_bytes, hipStream_t st) {
    return scan_run<uint32_t>(FirstOfRun{in, shift}, StoreFirst{in, out, shift}, true, n, count, ws, ws_bytes, st);
}

}  // namespace yx
