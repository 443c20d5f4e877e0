// Handle, stream, workspace and memory entry points of liboxiblas_mi355x.so.
#include "oxi_internal.hpp"

#include <hip/hip_ext.h>
#include <cstdlib>

#include <cstdarg>
#include <new>

int oxi_set_error(oxi_handle_t h, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->last_error = buf;
    return code;
}

// Grow-only device buffer: the old block is freed after the handle's streams
// are idle, the new one rounded up to 64 MiB granules so growth is rare.
static int grow(oxi_handle_t h, void*& buf, size_t& have, size_t bytes, void** out) {
    if (bytes <= have) { *out = buf; return OXI_OK; }
    size_t want = ((bytes + (size_t(64) << 20) - 1) >> 26) << 26;
    if (buf) {
        OXI_HIP(h, hipStreamSynchronize(h->stream));
        OXI_HIP(h, hipStreamSynchronize(h->side));
        if (h->bulk) OXI_HIP(h, hipStreamSynchronize(h->bulk));
        OXI_HIP(h, hipFree(buf));
        buf = nullptr; have = 0;
    }
    if (hipMalloc(&buf, want) != hipSuccess) {
        buf = nullptr;
        return oxi_set_error(h, OXI_ERR_ALLOC, "workspace allocation of %zu bytes failed", want);
    }
    have = want;
    *out = buf;
    return OXI_OK;
}

int oxi_workspace(oxi_handle_t h, size_t bytes, void** out) { return grow(h, h->ws, h->ws_bytes, bytes, out); }
int oxi_workspace2(oxi_handle_t h, size_t bytes, void** out) { return grow(h, h->ws2, h->ws2_bytes, bytes, out); }
int oxi_workspace3(oxi_handle_t h, size_t bytes, void** out) { return grow(h, h->ws3, h->ws3_bytes, bytes, out); }

extern "C" {

const char* oxi_version(void) { return "oxiblas_mi355x 0.1.0 (gfx950)"; }

int oxi_create(int device, oxi_handle_t* handle) {
    if (!handle) return -2;
    *handle = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return OXI_ERR_HIP;
    if (device < 0 || device >= count) return -1;
    oxi_handle_s* h = new (std::nothrow) oxi_handle_s();
    if (!h) return OXI_ERR_ALLOC;
    h->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete h; return OXI_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->num_cus = prop.multiProcessorCount;
    // default-priority streams only
    if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_a, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_b, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&h->dev_flags, 64 * sizeof(int)) != hipSuccess ||
        hipMalloc(&h->dev_info, 8 * sizeof(i64)) != hipSuccess ||
        hipMalloc(&h->trsm_inv, size_t(TRSM_INV_BLOCKS) * 4096 * sizeof(double)) != hipSuccess) {
        oxi_destroy(h);
        return OXI_ERR_HIP;
    }
    // The bulk stream's CU mask: bit i of the mask selects a CU of XCD i % 8
    // (measured on MI355X), so clearing the top `panel_cus` bits frees
    // panel_cus / 8 CUs on every XCD.  OXIBLAS_PANEL_CUS overrides the count
    // (a multiple of 8; 0 runs the trailing updates on the caller's stream).
    h->panel_cus = 32;
    if (const char* e = getenv("OXIBLAS_PANEL_CUS")) h->panel_cus = atoi(e);
    if (h->panel_cus < 0 || h->panel_cus % 8 || h->panel_cus > h->num_cus / 2) h->panel_cus = 0;
    if (h->panel_cus > 0) {
        std::vector<uint32_t> mask(size_t(h->num_cus + 31) / 32, 0u);
        for (int i = 0; i < h->num_cus - h->panel_cus; ++i) mask[i / 32] |= 1u << (i % 32);
        if (hipExtStreamCreateWithCUMask(&h->bulk, uint32_t(mask.size()), mask.data()) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_c, hipEventDisableTiming) != hipSuccess) {
            oxi_destroy(h);
            return OXI_ERR_HIP;
        }
    }
    (void)hipMemset(h->dev_flags, 0, 64 * sizeof(int));
    (void)hipMemset(h->dev_info, 0, 8 * sizeof(i64));
    (void)hipDeviceSynchronize();
    *handle = h;
    return OXI_OK;
}

int oxi_destroy(oxi_handle_t h) {
    if (!h) return OXI_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->side) (void)hipStreamSynchronize(h->side);
    if (h->bulk) (void)hipStreamSynchronize(h->bulk);
    (void)hipDeviceSynchronize();
    if (h->ws) (void)hipFree(h->ws);
    if (h->ws2) (void)hipFree(h->ws2);
    if (h->ws3) (void)hipFree(h->ws3);
    if (h->dev_flags) (void)hipFree(h->dev_flags);
    if (h->dev_info) (void)hipFree(h->dev_info);
    if (h->trsm_inv) (void)hipFree(h->trsm_inv);
    if (h->ev_a) (void)hipEventDestroy(h->ev_a);
    if (h->ev_b) (void)hipEventDestroy(h->ev_b);
    if (h->ev_c) (void)hipEventDestroy(h->ev_c);
    if (h->side) (void)hipStreamDestroy(h->side);
    if (h->bulk) (void)hipStreamDestroy(h->bulk);
    delete h;
    return OXI_OK;
}

int oxi_set_stream(oxi_handle_t h, void* stream) {
    if (!h) return -1;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    h->stream = (hipStream_t)stream;
    return OXI_OK;
}

int oxi_sync(oxi_handle_t h) {
    if (!h) return -1;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    (void)hipSetDevice(h->device);
    OXI_HIP(h, hipStreamSynchronize(h->stream));
    OXI_HIP(h, hipStreamSynchronize(h->side));
    if (h->bulk) OXI_HIP(h, hipStreamSynchronize(h->bulk));
    return OXI_OK;
}

const char* oxi_last_error(oxi_handle_t h) { return h ? h->last_error.c_str() : "null handle"; }

int oxi_malloc(oxi_handle_t h, void** ptr, uint64_t bytes) {
    if (!h) return -1;
    if (!ptr) return -2;
    (void)hipSetDevice(h->device);
    *ptr = nullptr;
    if (bytes == 0) return OXI_OK;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        *ptr = nullptr;
        return oxi_set_error(h, OXI_ERR_ALLOC, "hipMalloc(%llu) failed", (unsigned long long)bytes);
    }
    return OXI_OK;
}

int oxi_free(oxi_handle_t h, void* ptr) {
    if (!h) return -1;
    (void)hipSetDevice(h->device);
    if (ptr) OXI_HIP(h, hipFree(ptr));
    return OXI_OK;
}

int oxi_memcpy(oxi_handle_t h, void* dst, const void* src, uint64_t bytes, int kind) {
    if (!h) return -1;
    if (kind < 0 || kind > 2) return -5;
    if (bytes == 0) return OXI_OK;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    (void)hipSetDevice(h->device);
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    OXI_HIP(h, hipMemcpyAsync(dst, src, bytes, k, h->stream));
    OXI_HIP(h, hipStreamSynchronize(h->stream));
    return OXI_OK;
}

static int resolve_profile(oxi_handle_t h) {
    for (auto& e : h->prof_events) {
        OXI_HIP(h, hipEventSynchronize(e.b));
        float ms = 0;
        OXI_HIP(h, hipEventElapsedTime(&ms, e.a, e.b));
        h->prof_gemm_ms += ms;
        h->prof_gemm_flops += e.flops;
        h->prof_gemm_launches += 1;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    h->prof_events.clear();
    return OXI_OK;
}

int oxi_profile_enable(oxi_handle_t h, int enable) {
    if (!h) return -1;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    (void)resolve_profile(h);
    h->prof_enabled = enable;
    h->prof_gemm_ms = 0; h->prof_gemm_flops = 0; h->prof_gemm_launches = 0;
    return OXI_OK;
}

int oxi_profile_read(oxi_handle_t h, double* ms, double* flops, int64_t* launches) {
    if (!h) return -1;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    OXI_TRY(resolve_profile(h));
    if (ms) *ms = h->prof_gemm_ms;
    if (flops) *flops = h->prof_gemm_flops;
    if (launches) *launches = h->prof_gemm_launches;
    return OXI_OK;
}

}  // extern "C"

int oxi_prof_begin(oxi_handle_t h, hipStream_t s, hipEvent_t* a) {
    *a = nullptr;
    if (!h->prof_enabled) return OXI_OK;
    OXI_HIP(h, hipEventCreate(a));
    OXI_HIP(h, hipEventRecord(*a, s));
    return OXI_OK;
}

int oxi_prof_end(oxi_handle_t h, hipStream_t s, hipEvent_t a, double flops) {
    if (!h->prof_enabled || !a) return OXI_OK;
    hipEvent_t b;
    OXI_HIP(h, hipEventCreate(&b));
    OXI_HIP(h, hipEventRecord(b, s));
    h->prof_events.push_back({a, b, flops});
    return OXI_OK;
}
