// Runtime plumbing: per-thread error strings, device properties, fill kernel
// and the raw device-memory helpers of the C ABI.
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "internal.hpp"

namespace b40c {

namespace {
thread_local char g_last_error[1024] = "";
}

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

b40c_status fail(b40c_status code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

b40c_status device_info(DeviceInfo& info) {
    static std::mutex mu;
    static bool cached = false;
    static DeviceInfo cache;
    int dev = 0;
    B40C_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (!cached || cache.device != dev) {
        hipDeviceProp_t prop;
        B40C_HIP(hipGetDeviceProperties(&prop, dev));
        cache.device = dev;
        cache.num_cus = prop.multiProcessorCount;
        cache.arch = prop.gcnArchName;
        cached = true;
    }
    info = cache;
    return B40C_OK;
}

namespace {
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
}  // namespace

b40c_status fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t stream) {
    if (count == 0) return B40C_OK;
    size_t blocks = (count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, stream, p, value, count);
    B40C_HIP(hipGetLastError());
    return B40C_OK;
}

}  // namespace b40c

extern "C" {

const char* b40c_last_error_string(void) { return b40c::g_last_error; }

b40c_status b40c_device_info(int32_t* num_cus, char* arch, int32_t arch_len) {
    b40c::DeviceInfo info;
    B40C_TRY(b40c::device_info(info));
    if (num_cus) *num_cus = info.num_cus;
    if (arch && arch_len > 0) {
        std::strncpy(arch, info.arch.c_str(), (size_t)arch_len - 1);
        arch[arch_len - 1] = '\0';
    }
    return B40C_OK;
}

b40c_status b40c_device_alloc(void** d_ptr, int64_t bytes) {
    if (!d_ptr || bytes < 0) return b40c::fail(B40C_ERR_INVALID_ARG, "b40c_device_alloc: bad arguments");
    B40C_HIP(hipMalloc(d_ptr, bytes > 0 ? (size_t)bytes : 1));
    return B40C_OK;
}

b40c_status b40c_device_free(void* d_ptr) {
    if (d_ptr) B40C_HIP(hipFree(d_ptr));
    return B40C_OK;
}

b40c_status b40c_memcpy_h2d(void* d_dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!d_dst || !src))) return b40c::fail(B40C_ERR_INVALID_ARG, "b40c_memcpy_h2d");
    if (bytes) B40C_HIP(hipMemcpy(d_dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return B40C_OK;
}

b40c_status b40c_memcpy_d2h(void* dst, const void* d_src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !d_src))) return b40c::fail(B40C_ERR_INVALID_ARG, "b40c_memcpy_d2h");
    if (bytes) B40C_HIP(hipMemcpy(dst, d_src, (size_t)bytes, hipMemcpyDeviceToHost));
    return B40C_OK;
}

}  // extern "C"
