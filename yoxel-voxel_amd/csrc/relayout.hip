// Breadth-first re-layout of a VoxNode tree on the GPU: the traversal records
// and LOD colours of build_gpu_tree (scene.cpp), computed from the device copy
// of the node array that yx_renderer_update_scene keeps in sync page by page
// (CudaSVO::Update, reaction/report/main.tex:69-71).
//
// Level by level, one thread per node of the level:
//   count:  its present children (leaves + non-null node ids), with the id
//           checks of the host re-layout (id in range, no node reached twice);
//   scan:   exclusive sums give every node the first record of its children in
//           the next level and the position of its node children in the next
//           level's node list;
//   gather: the next level's node ids and their record indices.
// Then every level writes its node records and its leaves' records, and a
// bottom-up sweep sums leaf colours into the per-record mean colour. Integer
// work throughout: the result equals the host re-layout bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "internal.h"

#define RL_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace yx {

namespace {

constexpr uint32_t kNull = 0x80000000u;

struct RlFlags {
    uint32_t bad_id;  // child id >= number of nodes (the largest found)
    uint32_t shared;  // a node reached twice (the largest such id)
    uint32_t leaf;    // some node of the level has a leaf
    uint32_t any_bad;
};

__global__ void rl_count_kernel(const uint32_t* __restrict__ vox, uint32_t nn, const uint32_t* __restrict__ ids,
                                uint32_t n, uint32_t* __restrict__ nrec, uint32_t* __restrict__ nnode,
                                uint32_t* __restrict__ seen, RlFlags* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* v = vox + (size_t)ids[i] * 10;
    const uint32_t lm = v[0] & 0xffu;
    uint32_t cm = 0;
    for (int c = 0; c < 8; ++c) {
        if ((lm >> c) & 1u) continue;
        const uint32_t ch = v[2 + c];
        if (ch & kNull) continue;
        if (ch >= nn) {
            atomicOr(&flags->any_bad, 1u);
            atomicMax(&flags->bad_id, ch);
            continue;
        }
        if (atomicExch(&seen[ch], 1u) != 0u) {
            atomicOr(&flags->any_bad, 2u);
            atomicMax(&flags->shared, ch);
        }
        cm |= 1u << c;
    }
    nrec[i] = (uint32_t)__popc(lm | cm);
    nnode[i] = (uint32_t)__popc(cm);
    if (lm) atomicOr(&flags->leaf, 1u);
}

__global__ void rl_gather_kernel(const uint32_t* __restrict__ vox, const uint32_t* __restrict__ ids, uint32_t n,
                                 const uint32_t* __restrict__ roff, const uint32_t* __restrict__ noff,
                                 uint32_t begin_next, uint32_t* __restrict__ next_ids,
                                 uint32_t* __restrict__ next_rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* v = vox + (size_t)ids[i] * 10;
    const uint32_t lm = v[0] & 0xffu;
    const uint32_t first = begin_next + roff[i];
    uint32_t pos = noff[i], k = 0;
    for (int c = 0; c < 8; ++c) {
        if ((lm >> c) & 1u) {
            ++k;
            continue;
        }
        const uint32_t ch = v[2 + c];
        if (ch & kNull) continue;
        next_ids[pos] = ch;
        next_rec[pos] = first + k;
        ++pos;
        ++k;
    }
}

__global__ void rl_write_kernel(const uint32_t* __restrict__ vox, const uint32_t* __restrict__ ids,
                                const uint32_t* __restrict__ rec, uint32_t n, const uint32_t* __restrict__ roff,
                                uint32_t begin_next, uint2* __restrict__ recs, uint32_t* __restrict__ lod) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = ids[i];
    const uint32_t* v = vox + (size_t)id * 10;
    const uint32_t lm = v[0] & 0xffu;
    uint32_t cm = 0;
    for (int c = 0; c < 8; ++c)
        if (!((lm >> c) & 1u) && !(v[2 + c] & kNull)) cm |= 1u << c;
    if ((lm | cm) == 0u) {
        recs[rec[i]] = make_uint2(0u, 0u);
        return;
    }
    const uint32_t first = begin_next + roff[i];
    recs[rec[i]] = make_uint2(first, cm | (lm << 8));
    uint32_t k = 0;
    for (int c = 0; c < 8; ++c) {
        if ((lm >> c) & 1u) {
            recs[first + k] = make_uint2(v[2 + c], id);
            lod[first + k] = v[2 + c];
            ++k;
        } else if ((cm >> c) & 1u) {
            ++k;
        }
    }
}

// per node: channel sums and leaf count of its subtree (5 x u64)
__global__ void rl_lod_kernel(const uint32_t* __restrict__ vox, const uint32_t* __restrict__ ids,
                              const uint32_t* __restrict__ rec, uint32_t n, const uint32_t* __restrict__ noff,
                              const uint64_t* __restrict__ sums_next, uint64_t* __restrict__ sums,
                              uint32_t* __restrict__ lod) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* v = vox + (size_t)ids[i] * 10;
    const uint32_t lm = v[0] & 0xffu;
    uint64_t acc[5] = {0, 0, 0, 0, 0};
    uint32_t j = noff[i];
    for (int c = 0; c < 8; ++c) {
        const uint32_t ch = v[2 + c];
        if ((lm >> c) & 1u) {
            for (int k = 0; k < 4; ++k) acc[k] += (ch >> (8 * k)) & 0xffu;
            acc[4] += 1;
        } else if (!(ch & kNull)) {
            const uint64_t* s = sums_next + (size_t)j * 5;
            for (int k = 0; k < 5; ++k) acc[k] += s[k];
            ++j;
        }
    }
    for (int k = 0; k < 5; ++k) sums[(size_t)i * 5 + k] = acc[k];
    uint32_t mean = 0;
    if (acc[4] > 0)
        for (int k = 0; k < 4; ++k) mean |= (uint32_t)((acc[k] + acc[4] / 2) / acc[4]) << (8 * k);
    lod[rec[i]] = mean;
}

struct RlBufs {
    std::vector<void*> ptrs;
    ~RlBufs() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t n) {
        *p = nullptr;
        if (hipMalloc((void**)p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) {
            *p = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) in the GPU re-layout", n * sizeof(T));
        }
        ptrs.push_back(*p);
        return YX_OK;
    }
};

struct RlLevel {
    uint32_t n = 0;           // nodes of the level
    uint32_t* ids = nullptr;  // canonical ids
    uint32_t* rec = nullptr;  // their record indices
    uint32_t* roff = nullptr;
    uint32_t* noff = nullptr;
    uint32_t begin_next = 0;  // first record of the next level
};

inline dim3 rl_grid(uint32_t n) { return dim3((n + 255u) / 256u); }

}  // namespace

int gpu_relayout(const uint32_t* d_vox, uint32_t nn, uint32_t root, void* stream, GpuLayout& out) {
    const hipStream_t st = (hipStream_t)stream;
    out = GpuLayout();
    if (root & kNull) return YX_OK;  // EmptyNode / FullNode root: nothing to hit
    if (root >= nn) return fail(YX_ERR_FORMAT, "root id %u out of range (%u nodes)", root, nn);
    RlBufs B;
    int rc;
    uint32_t *d_seen = nullptr, *d_tmp_n = nullptr;
    RlFlags* d_flags = nullptr;
    if ((rc = B.alloc(&d_seen, nn)) != YX_OK || (rc = B.alloc(&d_flags, 1)) != YX_OK ||
        (rc = B.alloc(&d_tmp_n, 2)) != YX_OK)
        return rc;
    RL_HIP(hipMemsetAsync(d_seen, 0, (size_t)nn * sizeof(uint32_t), st));
    const uint32_t one = 1u, zero = 0u;
    RL_HIP(hipMemcpyAsync(d_seen + root, &one, sizeof(uint32_t), hipMemcpyHostToDevice, st));

    std::vector<RlLevel> lv(1);
    if ((rc = B.alloc(&lv[0].ids, 1)) != YX_OK || (rc = B.alloc(&lv[0].rec, 1)) != YX_OK) return rc;
    RL_HIP(hipMemcpyAsync(lv[0].ids, &root, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    RL_HIP(hipMemcpyAsync(lv[0].rec, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    lv[0].n = 1;

    uint64_t total = 1;  // records so far (the root's)
    int depth = 0;
    std::vector<uint32_t> level_records(1, 1u);
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    uint32_t* d_nrec = nullptr;
    uint32_t* d_nnode = nullptr;
    uint32_t cap_n = 0;
    for (size_t l = 0; lv[l].n > 0; ++l) {
        const uint32_t n = lv[l].n;
        if (l > 0) level_records.push_back((uint32_t)total);
        if (n > cap_n) {
            cap_n = std::max(n, cap_n * 2u);
            if ((rc = B.alloc(&d_nrec, cap_n)) != YX_OK || (rc = B.alloc(&d_nnode, cap_n)) != YX_OK) return rc;
            size_t bytes = 0;
            RL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, d_nrec, d_nrec, (int)cap_n, st));
            if (bytes > scan_bytes) {
                if ((rc = B.alloc((uint8_t**)&scan_tmp, bytes)) != YX_OK) return rc;
                scan_bytes = bytes;
            }
        }
        RL_HIP(hipMemsetAsync(d_flags, 0, sizeof(RlFlags), st));
        hipLaunchKernelGGL(rl_count_kernel, rl_grid(n), dim3(256), 0, st, d_vox, nn, lv[l].ids, n, d_nrec, d_nnode,
                           d_seen, d_flags);
        RL_HIP(hipGetLastError());
        if ((rc = B.alloc(&lv[l].roff, n)) != YX_OK || (rc = B.alloc(&lv[l].noff, n)) != YX_OK) return rc;
        RL_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, d_nrec, lv[l].roff, (int)n, st));
        RL_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, d_nnode, lv[l].noff, (int)n, st));
        RlFlags h_flags;
        uint32_t last[4];  // roff, nrec, noff, nnode of the last node
        RL_HIP(hipMemcpyAsync(&h_flags, d_flags, sizeof(RlFlags), hipMemcpyDeviceToHost, st));
        RL_HIP(hipMemcpyAsync(&last[0], lv[l].roff + (n - 1), 4, hipMemcpyDeviceToHost, st));
        RL_HIP(hipMemcpyAsync(&last[1], d_nrec + (n - 1), 4, hipMemcpyDeviceToHost, st));
        RL_HIP(hipMemcpyAsync(&last[2], lv[l].noff + (n - 1), 4, hipMemcpyDeviceToHost, st));
        RL_HIP(hipMemcpyAsync(&last[3], d_nnode + (n - 1), 4, hipMemcpyDeviceToHost, st));
        RL_HIP(hipStreamSynchronize(st));
        if (h_flags.any_bad & 1u) return fail(YX_ERR_FORMAT, "child id %u out of range (%u nodes)", h_flags.bad_id, nn);
        if (h_flags.any_bad & 2u) return fail(YX_ERR_FORMAT, "node %u reached twice (not a tree)", h_flags.shared);
        const uint64_t child_recs = (uint64_t)last[0] + last[1];
        const uint32_t child_nodes = last[2] + last[3];
        if (child_recs > 0 && (int)l + 1 > kMaxDepth) return fail(YX_ERR_RANGE, "tree deeper than %d levels", kMaxDepth);
        if (h_flags.leaf) depth = (int)l + 1;
        lv[l].begin_next = (uint32_t)total;
        total += child_recs;
        if (total + 8 > 0xFFFFFFF0ull) return fail(YX_ERR_RANGE, "more than 2^32 records");
        lv.emplace_back();
        RlLevel& nx = lv.back();
        nx.n = child_nodes;
        if (child_nodes > 0) {
            if ((rc = B.alloc(&nx.ids, child_nodes)) != YX_OK || (rc = B.alloc(&nx.rec, child_nodes)) != YX_OK) return rc;
            hipLaunchKernelGGL(rl_gather_kernel, rl_grid(n), dim3(256), 0, st, d_vox, lv[l].ids, n, lv[l].roff,
                               lv[l].noff, lv[l].begin_next, nx.ids, nx.rec);
            RL_HIP(hipGetLastError());
        }
    }
    const size_t levels = lv.size() - 1;  // levels with nodes
    level_records.push_back((uint32_t)total);
    if (depth == 0) return YX_OK;  // no leaf anywhere: nothing to hit

    uint2* d_recs = nullptr;
    uint32_t* d_lod = nullptr;
    if (hipMalloc((void**)&d_recs, total * sizeof(uint2)) != hipSuccess)
        return fail(YX_ERR_OOM, "hipMalloc(%zu) for the records", (size_t)(total * sizeof(uint2)));
    if (hipMalloc((void**)&d_lod, total * sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(d_recs);
        return fail(YX_ERR_OOM, "hipMalloc(%zu) for the LOD colours", (size_t)(total * sizeof(uint32_t)));
    }
    auto give_up = [&](int code) {
        (void)hipFree(d_recs);
        (void)hipFree(d_lod);
        return code;
    };
    for (size_t l = 0; l < levels; ++l) {
        hipLaunchKernelGGL(rl_write_kernel, rl_grid(lv[l].n), dim3(256), 0, st, d_vox, lv[l].ids, lv[l].rec, lv[l].n,
                           lv[l].roff, lv[l].begin_next, d_recs, d_lod);
        if (hipGetLastError() != hipSuccess) return give_up(fail(YX_ERR_HIP, "re-layout write kernel launch"));
    }
    // bottom-up subtree sums: two buffers of 5 u64 per node of the widest level
    uint32_t widest = 0;
    for (size_t l = 0; l < levels; ++l) widest = std::max(widest, lv[l].n);
    uint64_t *d_sa = nullptr, *d_sb = nullptr;
    if ((rc = B.alloc(&d_sa, (size_t)widest * 5)) != YX_OK || (rc = B.alloc(&d_sb, (size_t)widest * 5)) != YX_OK)
        return give_up(rc);
    for (size_t l = levels; l-- > 0;) {
        hipLaunchKernelGGL(rl_lod_kernel, rl_grid(lv[l].n), dim3(256), 0, st, d_vox, lv[l].ids, lv[l].rec, lv[l].n,
                           lv[l].noff, (const uint64_t*)d_sb, d_sa, d_lod);
        if (hipGetLastError() != hipSuccess) return give_up(fail(YX_ERR_HIP, "re-layout LOD kernel launch"));
        std::swap(d_sa, d_sb);
    }
    if (hipStreamSynchronize(st) != hipSuccess) return give_up(fail(YX_ERR_HIP, "re-layout"));
    out.recs = reinterpret_cast<uint32_t*>(d_recs);
    out.lod = d_lod;
    out.n_recs = (size_t)total;
    out.depth = depth;
    out.node_levels = (int)levels;
    out.level_records = level_records;
    out.empty = false;
    return YX_OK;
}

}  // namespace yx
