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

#include <algorithm>
#include <new>
#include <vector>

#include "internal.h"
#include "prims.h"

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

struct RlLevel {
    uint32_t n = 0;           // nodes of the level
    uint32_t* ids = nullptr;  // canonical ids
    uint32_t* rec = nullptr;  // their record indices
    uint32_t* roff = nullptr; // first child record, relative to begin_next
    uint32_t* noff = nullptr; // position of the first node child in the next level
    uint32_t begin_next = 0;  // first record of the next level
};

// one level's totals and the id checks, copied to the host in one transfer
struct RlTail {
    uint32_t any_bad, bad_id, shared, pad;
    uint64_t child_recs, child_nodes;
};

__global__ void rl_tail_kernel(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ nrec,
                               const uint32_t* __restrict__ noff, const uint32_t* __restrict__ nnode, uint32_t n,
                               RlFlags* flags, RlTail* tail) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    RlTail t;
    t.any_bad = flags->any_bad;
    t.bad_id = flags->bad_id;
    t.shared = flags->shared;
    t.pad = 0;
    t.child_recs = (uint64_t)roff[n - 1] + nrec[n - 1];
    t.child_nodes = (uint64_t)noff[n - 1] + nnode[n - 1];
    *tail = t;
    flags->any_bad = 0;
    flags->bad_id = 0;
    flags->shared = 0;
}

__global__ void rl_init_kernel(uint32_t* __restrict__ seen, uint32_t nn, uint32_t root, uint32_t* ids, uint32_t* rec,
                               RlFlags* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) seen[i] = (i == root) ? 1u : 0u;
    if (i == 0) {
        ids[0] = root;
        rec[0] = 0;
        flags->any_bad = 0;
        flags->bad_id = 0;
        flags->shared = 0;
    }
}

inline dim3 rl_grid(uint64_t n) { return dim3((unsigned)((n + 255u) / 256u)); }

template <class T>
int rl_grow(T** p, size_t* cap, size_t need, const char* what) {
    if (need <= *cap && *p) return YX_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc((void**)p, std::max<size_t>(need, 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return fail(YX_ERR_OOM, "hipMalloc(%zu) for %s", need * sizeof(T), what);
    }
    *cap = need;
    return YX_OK;
}

}  // namespace

// Device workspace of the re-layout, kept between calls: node-sized arrays
// (level lists laid end to end: level l at [off_l, off_l + n_l)), the subtree
// sums of the widest level, scan scratch and a pinned host copy of RlTail.
struct RelayoutWork {
    size_t cap = 0, cap_wide = 0, scan_bytes = 0;
    uint32_t *seen = nullptr, *ids = nullptr, *rec = nullptr, *roff = nullptr, *noff = nullptr;
    uint32_t *nrec = nullptr, *nnode = nullptr;
    uint64_t *sa = nullptr, *sb = nullptr;
    void* scan_tmp = nullptr;
    RlFlags* flags = nullptr;
    RlTail* tail = nullptr;
    RlTail* h_tail = nullptr;  // pinned
    ~RelayoutWork() {
        for (void* p : {(void*)seen, (void*)ids, (void*)rec, (void*)roff, (void*)noff, (void*)nrec, (void*)nnode,
                        (void*)sa, (void*)sb, scan_tmp, (void*)flags, (void*)tail})
            if (p) (void)hipFree(p);
        if (h_tail) (void)hipHostFree(h_tail);
    }
};

RelayoutWork* relayout_work_create() { return new (std::nothrow) RelayoutWork(); }
void relayout_work_free(RelayoutWork* w) { delete w; }

int gpu_relayout(const uint32_t* d_vox, uint32_t nn, uint32_t root, void* stream, RelayoutWork* w,
                 uint32_t** recs_io, uint32_t** lod_io, size_t* cap_io, GpuLayout& out) {
    const hipStream_t st = (hipStream_t)stream;
    out = GpuLayout();
    if (!w) return fail(YX_ERR_INVALID_ARG, "no re-layout workspace");
    if (root & kNull) return YX_OK;  // EmptyNode / FullNode root: nothing to hit
    if (root >= nn) return fail(YX_ERR_FORMAT, "root id %u out of range (%u nodes)", root, nn);
    if (nn >= (1u << 29)) return fail(YX_ERR_RANGE, "at most 2^29 nodes");
    int rc;
    if (w->cap < nn) {
        size_t c = 0;
        for (uint32_t** p : {&w->seen, &w->ids, &w->rec, &w->roff, &w->noff, &w->nrec, &w->nnode}) {
            c = w->cap;
            if ((rc = rl_grow(p, &c, nn, "the re-layout workspace")) != YX_OK) {
                w->cap = 0;
                return rc;
            }
        }
        w->cap = nn;
        if ((rc = rl_grow((uint8_t**)&w->scan_tmp, &w->scan_bytes, scan_workspace_bytes(nn), "scan scratch")) !=
            YX_OK)
            return rc;
    }
    if (!w->flags) {
        size_t c = 0;
        if ((rc = rl_grow(&w->flags, &c, 1, "flags")) != YX_OK) return rc;
        c = 0;
        if ((rc = rl_grow(&w->tail, &c, 1, "flags")) != YX_OK) return rc;
        if (hipHostMalloc((void**)&w->h_tail, sizeof(RlTail), hipHostMallocDefault) != hipSuccess) {
            w->h_tail = nullptr;
            return fail(YX_ERR_OOM, "pinned host memory for the re-layout");
        }
    }
    hipLaunchKernelGGL(rl_init_kernel, rl_grid(nn), dim3(256), 0, st, w->seen, nn, root, w->ids, w->rec, w->flags);
    RL_HIP(hipGetLastError());

    std::vector<RlLevel> lv(1);
    lv[0].n = 1;
    lv[0].ids = w->ids;
    lv[0].rec = w->rec;
    lv[0].roff = w->roff;
    lv[0].noff = w->noff;
    size_t off = 0;  // first slot of the current level in the node-sized arrays
    uint64_t total = 1;  // records so far (the root's)
    int depth = 0;
    uint32_t widest = 1;
    std::vector<uint32_t> level_records(1, 1u);
    for (size_t l = 0; lv[l].n > 0; ++l) {
        const uint32_t n = lv[l].n;
        widest = std::max(widest, n);
        if (l > 0) level_records.push_back((uint32_t)total);
        hipLaunchKernelGGL(rl_count_kernel, rl_grid(n), dim3(256), 0, st, d_vox, nn, lv[l].ids, n, w->nrec, w->nnode,
                           w->seen, w->flags);
        RL_HIP(hipGetLastError());
        if ((rc = scan_excl_u32(w->nrec, lv[l].roff, n, nullptr, w->scan_tmp, w->scan_bytes, st)) != YX_OK ||
            (rc = scan_excl_u32(w->nnode, lv[l].noff, n, nullptr, w->scan_tmp, w->scan_bytes, st)) != YX_OK)
            return rc;
        hipLaunchKernelGGL(rl_tail_kernel, dim3(1), dim3(64), 0, st, lv[l].roff, w->nrec, lv[l].noff, w->nnode, n,
                           w->flags, w->tail);
        RL_HIP(hipGetLastError());
        RL_HIP(hipMemcpyAsync(w->h_tail, w->tail, sizeof(RlTail), hipMemcpyDeviceToHost, st));
        RL_HIP(hipStreamSynchronize(st));
        const RlTail t = *w->h_tail;
        if (t.any_bad & 1u) return fail(YX_ERR_FORMAT, "child id %u out of range (%u nodes)", t.bad_id, nn);
        if (t.any_bad & 2u) return fail(YX_ERR_FORMAT, "node %u reached twice (not a tree)", t.shared);
        if (t.child_recs > 0 && (int)l + 1 > kMaxDepth)
            return fail(YX_ERR_RANGE, "tree deeper than %d levels", kMaxDepth);
        if (t.child_recs > t.child_nodes) depth = (int)l + 1;  // some child is a leaf
        lv[l].begin_next = (uint32_t)total;
        total += t.child_recs;
        if (total + 8 > 0xFFFFFFF0ull) return fail(YX_ERR_RANGE, "more than 2^32 records");
        off += n;
        lv.emplace_back();
        RlLevel& nx = lv.back();
        nx.n = (uint32_t)t.child_nodes;
        // without shared nodes every node appears once: the lists fit in nn
        nx.ids = w->ids + off;
        nx.rec = w->rec + off;
        nx.roff = w->roff + off;
        nx.noff = w->noff + off;
        if (nx.n > 0) {
            hipLaunchKernelGGL(rl_gather_kernel, rl_grid(n), dim3(256), 0, st, d_vox, lv[l].ids, n, lv[l].roff,
                               lv[l].noff, lv[l].begin_next, nx.ids, nx.rec);
            RL_HIP(hipGetLastError());
        }
    }
    const size_t levels = lv.size() - 1;  // levels with nodes
    level_records.push_back((uint32_t)total);
    if (depth == 0) return YX_OK;  // no leaf anywhere: nothing to hit

    // the caller's record arrays, grown only when too small
    if (*cap_io < total || !*recs_io || !*lod_io) {
        uint32_t *nr = nullptr, *nl = nullptr;
        if (hipMalloc((void**)&nr, total * 8u) != hipSuccess)
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for the records", (size_t)(total * 8u));
        if (hipMalloc((void**)&nl, total * 4u) != hipSuccess) {
            (void)hipFree(nr);
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for the LOD colours", (size_t)(total * 4u));
        }
        if (*recs_io) (void)hipFree(*recs_io);
        if (*lod_io) (void)hipFree(*lod_io);
        *recs_io = nr;
        *lod_io = nl;
        *cap_io = (size_t)total;
    }
    uint2* d_recs = reinterpret_cast<uint2*>(*recs_io);
    uint32_t* d_lod = *lod_io;
    for (size_t l = 0; l < levels; ++l) {
        hipLaunchKernelGGL(rl_write_kernel, rl_grid(lv[l].n), dim3(256), 0, st, d_vox, lv[l].ids, lv[l].rec, lv[l].n,
                           lv[l].roff, lv[l].begin_next, d_recs, d_lod);
        RL_HIP(hipGetLastError());
    }
    // bottom-up subtree sums: two buffers of 5 u64 per node of the widest level
    if (w->cap_wide < widest) {
        size_t c = w->cap_wide * 5;
        if ((rc = rl_grow(&w->sa, &c, (size_t)widest * 5, "subtree sums")) != YX_OK) return rc;
        c = w->cap_wide * 5;
        if ((rc = rl_grow(&w->sb, &c, (size_t)widest * 5, "subtree sums")) != YX_OK) {
            w->cap_wide = 0;
            return rc;
        }
        w->cap_wide = widest;
    }
    uint64_t *sa = w->sa, *sb = w->sb;
    for (size_t l = levels; l-- > 0;) {
        hipLaunchKernelGGL(rl_lod_kernel, rl_grid(lv[l].n), dim3(256), 0, st, d_vox, lv[l].ids, lv[l].rec, lv[l].n,
                           lv[l].noff, (const uint64_t*)sb, sa, d_lod);
        RL_HIP(hipGetLastError());
        std::swap(sa, sb);
    }
    RL_HIP(hipStreamSynchronize(st));
    out.recs = *recs_io;
    out.lod = *lod_io;
    out.n_recs = (size_t)total;
    out.depth = depth;
    out.node_levels = (int)levels;
    out.level_records = level_records;
    out.empty = false;
    return YX_OK;
}

}  // namespace yx
