// Brick pool manager for brick volumes (the compositing volume's brick table +
// per-slot parameters): CuBrickPoolManager::allocMap
// (reaction/voxtree/brickpool/BrickPool.cpp:21-62) and UpdateBrickPool
// (reaction/voxtree/density.cu:240-284), on the GPU.
//
// The pool manages arrays its caller owns (yx_volume_dev: the int32 brick
// table and the float4 parameters of `num_slots` slots). The reference
// allocates from the least-filled slices of a 3D texture and enumerates the
// mapped bricks in request order; here the pool is one linear HBM array and
// the policy is: the bricks of a request that have no slot, in order of their
// first appearance in the list, get the lowest free slots in increasing order.
// Allocation runs as claim (atomicMin of the list position per brick) -> flag
// -> scan (rank of each new brick) -> free-slot enumeration (scan over the
// used map) -> assign, so the result does not depend on thread timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>

#include "internal.h"
#include "prims.h"

#define BP_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace yx {
namespace {

struct BpFlags {
    uint32_t bad;       // 1: brick index out of range; 2: a slot referenced twice / out of range
    uint32_t released;  // slots freed by a release
};

__global__ void bp_init_used_kernel(const int32_t* __restrict__ table, uint32_t nb, uint32_t cap,
                                    uint32_t* __restrict__ used, BpFlags* flags) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int32_t s = table[b];
    if (s < 0) return;
    if ((uint32_t)s >= cap || atomicExch(&used[s], 1u) != 0u) atomicOr(&flags->bad, 2u);
}

__global__ void bp_fill_kernel(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void bp_claim_kernel(const int32_t* __restrict__ bricks, uint32_t n, uint32_t nb,
                                uint32_t* __restrict__ claim, BpFlags* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bricks[i];
    if (b < 0 || (uint32_t)b >= nb) {
        atomicOr(&flags->bad, 1u);
        return;
    }
    atomicMin(&claim[b], i);
}

__global__ void bp_flag_kernel(const int32_t* __restrict__ bricks, uint32_t n, const uint32_t* __restrict__ claim,
                               const int32_t* __restrict__ table, uint32_t* __restrict__ flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bricks[i];
    flag[i] = (claim[b] == i && table[b] < 0) ? 1u : 0u;
}

__global__ void bp_free_flag_kernel(const uint32_t* __restrict__ used, uint32_t cap, uint32_t* __restrict__ isfree) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < cap) isfree[j] = used[j] ? 0u : 1u;
}

__global__ void bp_free_list_kernel(const uint32_t* __restrict__ isfree, const uint32_t* __restrict__ fpos,
                                    uint32_t cap, uint32_t* __restrict__ freelist) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < cap && isfree[j]) freelist[fpos[j]] = j;
}

__global__ void bp_assign_kernel(const int32_t* __restrict__ bricks, uint32_t n, const uint32_t* __restrict__ flag,
                                 const uint32_t* __restrict__ rank, const uint32_t* __restrict__ freelist,
                                 int32_t* __restrict__ table, uint32_t* __restrict__ used) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t s = freelist[rank[i]];
    table[bricks[i]] = (int32_t)s;
    used[s] = 1u;
}

__global__ void bp_finish_kernel(const int32_t* __restrict__ bricks, uint32_t n, const int32_t* __restrict__ table,
                                 uint32_t* __restrict__ claim, int32_t* __restrict__ slots_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bricks[i];
    if (slots_out) slots_out[i] = table[b];
    claim[b] = 0xFFFFFFFFu;
}

__global__ void bp_release_kernel(const int32_t* __restrict__ bricks, uint32_t n, uint32_t nb,
                                  int32_t* __restrict__ table, uint32_t* __restrict__ used, BpFlags* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bricks[i];
    if (b < 0 || (uint32_t)b >= nb) {
        atomicOr(&flags->bad, 1u);
        return;
    }
    const int32_t s = atomicExch(&table[b], -1);
    if (s >= 0) {
        used[s] = 0u;
        atomicAdd(&flags->released, 1u);
    }
}

// a brick listed more than once is written by its last occurrence: claim the
// smallest n - 1 - i
__global__ void bp_claim_last_kernel(const int32_t* __restrict__ bricks, uint32_t n, uint32_t nb,
                                     uint32_t* __restrict__ claim, BpFlags* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bricks[i];
    if (b < 0 || (uint32_t)b >= nb) {
        atomicOr(&flags->bad, 1u);
        return;
    }
    atomicMin(&claim[b], n - 1u - i);
}

// one workgroup per listed brick: its packed voxels into its slot
__global__ __launch_bounds__(256) void bp_update_kernel(const int32_t* __restrict__ bricks, uint32_t n,
                                                        const int32_t* __restrict__ table,
                                                        const uint32_t* __restrict__ claim, uint32_t vox,
                                                        const float4* __restrict__ packed,
                                                        float4* __restrict__ params) {
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const int32_t b = bricks[i];
    if (claim[b] != n - 1u - i) return;
    const int32_t s = table[b];
    if (s < 0) return;
    const float4* src = packed + (size_t)i * vox;
    float4* dst = params + (size_t)s * vox;
    for (uint32_t t = threadIdx.x; t < vox; t += blockDim.x) dst[t] = src[t];
}

inline dim3 bp_grid(uint64_t n) { return dim3((unsigned)std::max<uint64_t>(1, (n + 255u) / 256u)); }

}  // namespace
}  // namespace yx

using namespace yx;

struct yx_brick_pool {
    yx_volume_dev vol;  // caller-owned table and params
    uint32_t nb = 0;    // bricks in the table: (grid / brick)^3
    uint32_t vox = 0;   // voxels per brick
    uint32_t used_count = 0;
    uint32_t* d_used = nullptr;   // [num_slots]
    uint32_t* d_claim = nullptr;  // [nb], all ones between calls
    BpFlags* d_flags = nullptr;
    uint32_t *d_flag = nullptr, *d_rank = nullptr;  // [cap_list]
    uint32_t *d_isfree = nullptr, *d_fpos = nullptr, *d_freelist = nullptr;  // [num_slots]
    uint32_t cap_list = 0;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    ~yx_brick_pool() {
        for (void* p : {(void*)d_used, (void*)d_claim, (void*)d_flags, (void*)d_flag, (void*)d_rank, (void*)d_isfree,
                        (void*)d_fpos, (void*)d_freelist, scan_tmp})
            if (p) (void)hipFree(p);
    }
};

template <class T>
static int bp_alloc(T** p, size_t n) {
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return fail(YX_ERR_OOM, "hipMalloc(%zu) for the brick pool", n * sizeof(T));
    }
    return YX_OK;
}

static int bp_scan_space(yx_brick_pool* p, uint32_t n, hipStream_t st) {
    (void)st;
    const size_t bytes = scan_workspace_bytes(std::max(n, 1u));
    if (bytes <= p->scan_bytes) return YX_OK;
    if (p->scan_tmp) BP_HIP(hipFree(p->scan_tmp));
    p->scan_tmp = nullptr;
    p->scan_bytes = 0;
    int rc = bp_alloc((uint8_t**)&p->scan_tmp, bytes);
    if (rc != YX_OK) return rc;
    p->scan_bytes = bytes;
    return YX_OK;
}

extern "C" int yx_brick_pool_create(const yx_volume_dev* vol, void* stream, yx_brick_pool** out) {
    clear_error();
    if (!vol || !out || !vol->table || (!vol->params && vol->num_slots))
        return fail(YX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (vol->brick < 1 || vol->grid < vol->brick || vol->grid % vol->brick)
        return fail(YX_ERR_INVALID_ARG, "grid %d must be a positive multiple of brick %d", vol->grid, vol->brick);
    const uint64_t g = (uint64_t)(vol->grid / vol->brick);
    if (g * g * g >= 0x80000000ull || (uint64_t)vol->brick * vol->brick * vol->brick > (1u << 20))
        return fail(YX_ERR_RANGE, "volume too large for the brick pool");
    if (vol->num_slots >= 0x80000000u) return fail(YX_ERR_RANGE, "at most 2^31 slots");
    const hipStream_t st = (hipStream_t)stream;
    yx_brick_pool* p = new (std::nothrow) yx_brick_pool();
    if (!p) return fail(YX_ERR_OOM, "out of host memory");
    p->vol = *vol;
    p->nb = (uint32_t)(g * g * g);
    p->vox = (uint32_t)(vol->brick * vol->brick * vol->brick);
    int rc;
    if ((rc = bp_alloc(&p->d_used, vol->num_slots)) != YX_OK || (rc = bp_alloc(&p->d_claim, p->nb)) != YX_OK ||
        (rc = bp_alloc(&p->d_flags, 1)) != YX_OK || (rc = bp_alloc(&p->d_isfree, vol->num_slots)) != YX_OK ||
        (rc = bp_alloc(&p->d_fpos, vol->num_slots)) != YX_OK ||
        (rc = bp_alloc(&p->d_freelist, vol->num_slots)) != YX_OK || (rc = bp_scan_space(p, vol->num_slots, st)) != YX_OK) {
        delete p;
        return rc;
    }
    auto give_up = [&](int code) {
        delete p;
        return code;
    };
    hipLaunchKernelGGL(bp_fill_kernel, bp_grid(vol->num_slots), dim3(256), 0, st, p->d_used, vol->num_slots, 0u);
    hipLaunchKernelGGL(bp_fill_kernel, bp_grid(p->nb), dim3(256), 0, st, p->d_claim, p->nb, 0xFFFFFFFFu);
    hipLaunchKernelGGL(bp_fill_kernel, dim3(1), dim3(64), 0, st, (uint32_t*)p->d_flags, 2u, 0u);
    hipLaunchKernelGGL(bp_init_used_kernel, bp_grid(p->nb), dim3(256), 0, st, vol->table, p->nb, vol->num_slots,
                       p->d_used, p->d_flags);
    if (hipGetLastError() != hipSuccess) return give_up(fail(YX_ERR_HIP, "brick pool set-up launch"));
    BpFlags f;
    if (hipMemcpyAsync(&f, p->d_flags, sizeof(f), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return give_up(fail(YX_ERR_HIP, "brick pool set-up"));
    if (f.bad) return give_up(fail(YX_ERR_FORMAT, "brick table references a slot twice or beyond %u slots",
                                   vol->num_slots));
    // used slots: count them with a reduction over the used map
    uint32_t* d_cnt = nullptr;
    if ((rc = bp_alloc(&d_cnt, 1)) != YX_OK) return give_up(rc);
    uint32_t cnt = 0;
    const bool ok = vol->num_slots == 0 ||
                    (reduce_sum_u32(p->d_used, vol->num_slots, d_cnt, p->scan_tmp, p->scan_bytes, st) == YX_OK &&
                     hipMemcpyAsync(&cnt, d_cnt, 4, hipMemcpyDeviceToHost, st) == hipSuccess &&
                     hipStreamSynchronize(st) == hipSuccess);
    (void)hipFree(d_cnt);
    if (!ok) return give_up(fail(YX_ERR_HIP, "counting used slots"));
    p->used_count = cnt;
    *out = p;
    return YX_OK;
}

extern "C" void yx_brick_pool_destroy(yx_brick_pool* p) { delete p; }

static int ensure_list(yx_brick_pool* p, uint32_t n, hipStream_t st) {
    if (n <= p->cap_list && p->d_flag) return YX_OK;
    if (p->d_flag) BP_HIP(hipFree(p->d_flag));
    if (p->d_rank) BP_HIP(hipFree(p->d_rank));
    p->d_flag = p->d_rank = nullptr;
    p->cap_list = 0;
    int rc;
    if ((rc = bp_alloc(&p->d_flag, n)) != YX_OK || (rc = bp_alloc(&p->d_rank, n)) != YX_OK) return rc;
    p->cap_list = n;
    return bp_scan_space(p, std::max(n, p->vol.num_slots), st);
}

extern "C" int yx_brick_pool_alloc(yx_brick_pool* p, const int32_t* bricks, uint32_t n, int32_t* slots_out,
                                   uint32_t* new_count, void* stream) {
    clear_error();
    if (!p || (!bricks && n)) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (new_count) *new_count = 0;
    if (n == 0) return YX_OK;
    if (n >= 0x80000000u) return fail(YX_ERR_RANGE, "at most 2^31 bricks per call");
    const hipStream_t st = (hipStream_t)stream;
    int rc = ensure_list(p, n, st);
    if (rc != YX_OK) return rc;
    const uint32_t cap = p->vol.num_slots;
    int32_t* table = const_cast<int32_t*>(p->vol.table);
    hipLaunchKernelGGL(bp_fill_kernel, dim3(1), dim3(64), 0, st, (uint32_t*)p->d_flags, 2u, 0u);
    hipLaunchKernelGGL(bp_claim_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, p->nb, p->d_claim, p->d_flags);
    BP_HIP(hipGetLastError());
    BpFlags f;
    BP_HIP(hipMemcpyAsync(&f, p->d_flags, sizeof(f), hipMemcpyDeviceToHost, st));
    BP_HIP(hipStreamSynchronize(st));
    if (f.bad) {
        // claims of the valid entries are undone before failing
        hipLaunchKernelGGL(bp_fill_kernel, bp_grid(p->nb), dim3(256), 0, st, p->d_claim, p->nb, 0xFFFFFFFFu);
        BP_HIP(hipStreamSynchronize(st));
        return fail(YX_ERR_RANGE, "brick index outside the %u bricks of the table", p->nb);
    }
    hipLaunchKernelGGL(bp_flag_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, p->d_claim, table, p->d_flag);
    if ((rc = scan_excl_u32(p->d_flag, p->d_rank, n, nullptr, p->scan_tmp, p->scan_bytes, st)) != YX_OK) return rc;
    // free slots in increasing order
    uint32_t last[2] = {0, 0}, free_last[2] = {0, 0};
    if (cap) {
        hipLaunchKernelGGL(bp_free_flag_kernel, bp_grid(cap), dim3(256), 0, st, p->d_used, cap, p->d_isfree);
        if ((rc = scan_excl_u32(p->d_isfree, p->d_fpos, cap, nullptr, p->scan_tmp, p->scan_bytes, st)) != YX_OK)
            return rc;
        hipLaunchKernelGGL(bp_free_list_kernel, bp_grid(cap), dim3(256), 0, st, p->d_isfree, p->d_fpos, cap,
                           p->d_freelist);
        BP_HIP(hipMemcpyAsync(&free_last[0], p->d_fpos + (cap - 1), 4, hipMemcpyDeviceToHost, st));
        BP_HIP(hipMemcpyAsync(&free_last[1], p->d_isfree + (cap - 1), 4, hipMemcpyDeviceToHost, st));
    }
    BP_HIP(hipMemcpyAsync(&last[0], p->d_rank + (n - 1), 4, hipMemcpyDeviceToHost, st));
    BP_HIP(hipMemcpyAsync(&last[1], p->d_flag + (n - 1), 4, hipMemcpyDeviceToHost, st));
    BP_HIP(hipStreamSynchronize(st));
    const uint32_t want = last[0] + last[1], have = free_last[0] + free_last[1];
    if (want > have) {
        hipLaunchKernelGGL(bp_finish_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, table, p->d_claim,
                           (int32_t*)nullptr);
        BP_HIP(hipStreamSynchronize(st));
        return fail(YX_ERR_RANGE, "%u new bricks but %u free slots (nothing allocated)", want, have);
    }
    hipLaunchKernelGGL(bp_assign_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, p->d_flag, p->d_rank,
                       p->d_freelist, table, p->d_used);
    hipLaunchKernelGGL(bp_finish_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, table, p->d_claim, slots_out);
    BP_HIP(hipGetLastError());
    p->used_count += want;
    if (new_count) *new_count = want;
    return YX_OK;
}

extern "C" int yx_brick_pool_release(yx_brick_pool* p, const int32_t* bricks, uint32_t n, uint32_t* released,
                                     void* stream) {
    clear_error();
    if (!p || (!bricks && n)) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (released) *released = 0;
    if (n == 0) return YX_OK;
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bp_fill_kernel, dim3(1), dim3(64), 0, st, (uint32_t*)p->d_flags, 2u, 0u);
    hipLaunchKernelGGL(bp_release_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, p->nb,
                       const_cast<int32_t*>(p->vol.table), p->d_used, p->d_flags);
    BP_HIP(hipGetLastError());
    BpFlags f;
    BP_HIP(hipMemcpyAsync(&f, p->d_flags, sizeof(f), hipMemcpyDeviceToHost, st));
    BP_HIP(hipStreamSynchronize(st));
    p->used_count -= f.released;
    if (released) *released = f.released;
    if (f.bad) return fail(YX_ERR_RANGE, "brick index outside the %u bricks of the table (valid ones released)", p->nb);
    return YX_OK;
}

extern "C" int yx_brick_pool_update(yx_brick_pool* p, const int32_t* bricks, uint32_t n, const float* packed,
                                    void* stream) {
    clear_error();
    if (!p || (n && (!bricks || !packed))) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (n == 0) return YX_OK;
    if (!p->vol.params) return fail(YX_ERR_INVALID_ARG, "the volume has no parameter slots");
    if (n >= 0x80000000u) return fail(YX_ERR_RANGE, "at most 2^31 bricks per call");
    const hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bp_fill_kernel, dim3(1), dim3(64), 0, st, (uint32_t*)p->d_flags, 2u, 0u);
    hipLaunchKernelGGL(bp_claim_last_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, p->nb, p->d_claim,
                       p->d_flags);
    BP_HIP(hipGetLastError());
    BpFlags f;
    BP_HIP(hipMemcpyAsync(&f, p->d_flags, sizeof(f), hipMemcpyDeviceToHost, st));
    BP_HIP(hipStreamSynchronize(st));
    if (f.bad) {
        hipLaunchKernelGGL(bp_fill_kernel, bp_grid(p->nb), dim3(256), 0, st, p->d_claim, p->nb, 0xFFFFFFFFu);
        BP_HIP(hipStreamSynchronize(st));
        return fail(YX_ERR_RANGE, "brick index outside the %u bricks of the table (nothing copied)", p->nb);
    }
    hipLaunchKernelGGL(bp_update_kernel, dim3(n), dim3(256), 0, st, bricks, n, p->vol.table, p->d_claim, p->vox,
                       reinterpret_cast<const float4*>(packed),
                       reinterpret_cast<float4*>(const_cast<float*>(p->vol.params)));
    hipLaunchKernelGGL(bp_finish_kernel, bp_grid(n), dim3(256), 0, st, bricks, n, p->vol.table, p->d_claim,
                       (int32_t*)nullptr);
    BP_HIP(hipGetLastError());
    return YX_OK;
}

extern "C" int yx_brick_pool_stats(const yx_brick_pool* p, uint32_t* used, uint32_t* capacity) {
    clear_error();
    if (!p || !used || !capacity) return fail(YX_ERR_INVALID_ARG, "null argument");
    *used = p->used_count;
    *capacity = p->vol.num_slots;
    return YX_OK;
}
