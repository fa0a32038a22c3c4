// Device-wide scans, a stable LSD radix sort and a select-distinct pass
// (gfx950), shared by the scene builders, the re-layout, the brick pool and
// the compositor's adjoint. All run on the caller's stream, take their
// workspace from the caller and never wait on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace yx {

// ---- device helpers (also used inside other kernels) ----

// Exclusive scan of one value per thread over a block of NT threads (NT a
// multiple of 64); `shw` holds NT / 64 words of LDS; `total` gets the block's
// sum. Contains two barriers: every thread of the block must call it.
template <class T, int NT>
__device__ __forceinline__ T block_excl_scan(T v, T* shw, T& total) {
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    T inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T o = __shfl_up(inc, (unsigned)off);
        if (lane >= off) inc += o;
    }
    if (lane == 63) shw[w] = inc;
    __syncthreads();
    T before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) {
        const T x = shw[k];
        before += k < w ? x : (T)0;
        tot += x;
    }
    total = tot;
    __syncthreads();
    return before + inc - v;
}

// Lanes of the wave holding the same `bits`-bit digit as this lane, among the
// lanes with `in` set (one ballot per bit).
__device__ __forceinline__ uint64_t digit_peers(bool in, uint32_t d, int bits) {
    uint64_t peers = __ballot(in);
    for (int b = 0; b < bits; ++b) {
        const bool on = (d >> b) & 1u;
        const uint64_t m = __ballot(on);
        peers &= on ? m : ~m;
    }
    return peers;
}

// ---- device-wide operations (host functions) ----

// Items per scan block: 256 threads x 16 coalesced rounds.
constexpr uint32_t kPrimScanChunk = 4096;
// Keys per sort tile (256 threads x 16 rounds); a sort block takes a run of
// consecutive tiles.
constexpr uint32_t kPrimSortTile = 4096;
constexpr int kPrimRadixBits = 8;
constexpr int kPrimRadix = 1 << kPrimRadixBits;

// Workspace bytes of the scans below for n items (partial sums of 8 bytes).
size_t scan_workspace_bytes(uint64_t n);

// out[i] = in[0] + ... + in[i - 1] for i < n (exclusive); out may alias in.
// *total (device memory, may be null) receives the sum of all n items.
// n < 2^32; the u32 sums wrap like unsigned arithmetic.
int scan_excl_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* ws, size_t ws_bytes,
                  hipStream_t st);
// The same over 0/1 (or any byte) flags, u32 sums.
int scan_excl_u8(const uint8_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* ws, size_t ws_bytes,
                 hipStream_t st);
// u32 items, u64 sums.
int scan_excl_u32_u64(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total, void* ws, size_t ws_bytes,
                      hipStream_t st);

// Sum of n u32 items into *total (device), u32 wrap.
int reduce_sum_u32(const uint32_t* in, uint64_t n, uint32_t* total, void* ws, size_t ws_bytes, hipStream_t st);

// Stable LSD radix sort by bits [b0, b1) of the keys, 8-bit digits, one
// histogram / scan / scatter per digit.
//   keys_in   the n keys (not written);
//   ids_in    their ids, or null: ids are the positions 0..n-1;
//   kbuf[2], ibuf[2]   ping-pong buffers of n keys / ids each (ibuf null:
//             keys only); the result is in one of them, returned in
//             *kout / *iout;
//   n_host / n_dev     the count: *n_dev (device) when n_dev is not null,
//             read by the kernels, with n_host as its upper bound (the
//             grids follow n_host); else n_host.
// Workspace: sort_workspace_bytes(n_host). n_host < 2^31.
size_t sort_workspace_bytes(uint64_t n_host);
template <class K>
int radix_sort(const K* keys_in, const uint32_t* ids_in, K* const kbuf[2], uint32_t* const ibuf[2], uint64_t n_host,
               const uint32_t* n_dev, int b0, int b1, const K** kout, const uint32_t** iout, void* ws,
               size_t ws_bytes, hipStream_t st);

// Distinct values of (in[i] >> shift) over a sorted array, in order, into
// out; *count (device) receives their number. n < 2^32.
size_t distinct_workspace_bytes(uint64_t n);
int distinct_shifted_u64(const uint64_t* in, uint64_t* out, uint64_t n, int shift, uint64_t* count, void* ws,
                         size_t ws_bytes, hipStream_t st);

}  // namespace yx
