// C entry points for the device primitives of prims.hip (scans, the stable
// LSD radix sort, select-distinct), so that they can be checked on their own
// against a host sort and scan. The library's builders call the primitives
// directly; these wrappers add only a workspace allocation and a wait.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"
#include "prims.h"

using namespace yx;

namespace {

// Device buffers freed on every exit path.
struct Scratch {
    void* p[5] = {};
    int n = 0;
    ~Scratch() {
        for (int i = 0; i < n; ++i) (void)hipFree(p[i]);
    }
    template <class T>
    int get(T** out, size_t bytes) {
        *out = nullptr;
        if (bytes == 0) bytes = 256;
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(YX_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        p[n++] = q;
        *out = (T*)q;
        return YX_OK;
    }
};

int finish(hipStream_t st) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(YX_ERR_HIP, "primitive: %s", hipGetErrorString(e));
    return YX_OK;
}

template <class K>
int sort_entry(const K* keys, const uint32_t* ids, uint64_t n, const uint32_t* n_dev, int32_t begin_bit,
               int32_t end_bit, K* keys_out, uint32_t* ids_out, void* stream) {
    clear_error();
    if ((n && !keys) || (n && !keys_out)) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (begin_bit < 0 || end_bit < begin_bit || end_bit > (int32_t)(8 * sizeof(K)))
        return fail(YX_ERR_INVALID_ARG, "sort bits [%d, %d)", begin_bit, end_bit);
    if (n >= 0x80000000ull) return fail(YX_ERR_RANGE, "sort of %llu keys (at most 2^31 - 1)", (unsigned long long)n);
    if (n == 0) return YX_OK;
    const hipStream_t st = (hipStream_t)stream;
    Scratch s;
    K* kb[2] = {};
    uint32_t* ib[2] = {};
    void* ws = nullptr;
    const size_t ws_bytes = sort_workspace_bytes(n);
    int rc;
    if ((rc = s.get(&kb[0], n * sizeof(K))) != YX_OK || (rc = s.get(&kb[1], n * sizeof(K))) != YX_OK ||
        (rc = s.get(&ws, ws_bytes)) != YX_OK)
        return rc;
    if (ids_out && ((rc = s.get(&ib[0], n * 4)) != YX_OK || (rc = s.get(&ib[1], n * 4)) != YX_OK)) return rc;
    K* const kbuf[2] = {kb[0], kb[1]};
    uint32_t* const ibuf[2] = {ib[0], ib[1]};
    const K* kres = nullptr;
    const uint32_t* ires = nullptr;
    rc = radix_sort<K>(keys, ids, kbuf, ids_out ? ibuf : nullptr, n, n_dev, begin_bit, end_bit, &kres, &ires, ws,
                       ws_bytes, st);
    if (rc != YX_OK) return rc;
    // an empty bit range leaves the input where it is (ids null: positions)
    if (hipMemcpyAsync(keys_out, kres, n * sizeof(K), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(YX_ERR_HIP, "sort result copy");
    if (ids_out) {
        if (ires) {
            if (hipMemcpyAsync(ids_out, ires, n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
                return fail(YX_ERR_HIP, "sort id copy");
        } else {
            return fail(YX_ERR_INVALID_ARG, "ids_out needs ids or a non-empty bit range");
        }
    }
    return finish(st);
}

}  // namespace

extern "C" int yx_prim_sort_u32(const uint32_t* keys, const uint32_t* ids, uint64_t n, const uint32_t* n_dev,
                                int32_t begin_bit, int32_t end_bit, uint32_t* keys_out, uint32_t* ids_out,
                                void* stream) {
    return sort_entry<uint32_t>(keys, ids, n, n_dev, begin_bit, end_bit, keys_out, ids_out, stream);
}

extern "C" int yx_prim_sort_u64(const uint64_t* keys, const uint32_t* ids, uint64_t n, const uint32_t* n_dev,
                                int32_t begin_bit, int32_t end_bit, uint64_t* keys_out, uint32_t* ids_out,
                                void* stream) {
    return sort_entry<uint64_t>(keys, ids, n, n_dev, begin_bit, end_bit, keys_out, ids_out, stream);
}

extern "C" int yx_prim_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* stream) {
    clear_error();
    if (n && (!in || !out)) return fail(YX_ERR_INVALID_ARG, "null argument");
    Scratch s;
    void* ws = nullptr;
    int rc = s.get(&ws, scan_workspace_bytes(n));
    if (rc != YX_OK) return rc;
    rc = scan_excl_u32(in, out, n, total, ws, scan_workspace_bytes(n), (hipStream_t)stream);
    return rc != YX_OK ? rc : finish((hipStream_t)stream);
}

extern "C" int yx_prim_scan_u8(const uint8_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* stream) {
    clear_error();
    if (n && (!in || !out)) return fail(YX_ERR_INVALID_ARG, "null argument");
    Scratch s;
    void* ws = nullptr;
    int rc = s.get(&ws, scan_workspace_bytes(n));
    if (rc != YX_OK) return rc;
    rc = scan_excl_u8(in, out, n, total, ws, scan_workspace_bytes(n), (hipStream_t)stream);
    return rc != YX_OK ? rc : finish((hipStream_t)stream);
}

extern "C" int yx_prim_scan_u32_u64(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total, void* stream) {
    clear_error();
    if (n && (!in || !out)) return fail(YX_ERR_INVALID_ARG, "null argument");
    Scratch s;
    void* ws = nullptr;
    int rc = s.get(&ws, scan_workspace_bytes(n));
    if (rc != YX_OK) return rc;
    rc = scan_excl_u32_u64(in, out, n, total, ws, scan_workspace_bytes(n), (hipStream_t)stream);
    return rc != YX_OK ? rc : finish((hipStream_t)stream);
}

extern "C" int yx_prim_reduce_u32(const uint32_t* in, uint64_t n, uint32_t* total, void* stream) {
    clear_error();
    if (!total || (n && !in)) return fail(YX_ERR_INVALID_ARG, "null argument");
    Scratch s;
    void* ws = nullptr;
    int rc = s.get(&ws, scan_workspace_bytes(n));
    if (rc != YX_OK) return rc;
    rc = reduce_sum_u32(in, n, total, ws, scan_workspace_bytes(n), (hipStream_t)stream);
    return rc != YX_OK ? rc : finish((hipStream_t)stream);
}

extern "C" int yx_prim_distinct_u64(const uint64_t* in, uint64_t* out, uint64_t n, int32_t shift, uint64_t* count,
                                    void* stream) {
    clear_error();
    if (!count || (n && (!in || !out))) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (shift < 0 || shift > 63) return fail(YX_ERR_INVALID_ARG, "shift %d", shift);
    Scratch s;
    void* ws = nullptr;
    int rc = s.get(&ws, distinct_workspace_bytes(n));
    if (rc != YX_OK) return rc;
    rc = distinct_shifted_u64(in, out, n, shift, count, ws, distinct_workspace_bytes(n), (hipStream_t)stream);
    return rc != YX_OK ? rc : finish((hipStream_t)stream);
}
