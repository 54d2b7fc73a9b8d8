// ImageStack .tmp image files (src/FileTMP.cpp:75-233) in and out of device
// memory.  A file is a header of five int32 (width, height, frames,
// channels, type code) followed by the samples in channel, frame, row,
// column order: the planar layout the library computes in, so a file maps
// onto a device buffer with no reordering.  Samples cross PCIe in the file's
// own type (a quarter of the bytes for 8-bit files) and are converted on the
// GPU; reading streams the file through two pinned buffers so the disk
// read, the copy and the conversion overlap.
//
// Conversions follow the reference's C casts as the x86-64 reference build
// evaluates them (src/FileTMP.cpp:43-56, 129-142):
//   load : (float)sample, rounded to nearest (exact below 2^24);
//   save : (T)value truncates toward zero; for 8-, 16- and 32-bit signed
//          types through int32 (outside int32: INT32_MIN), for uint32
//          through int64, keeping the low bits; float64 widens exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gt_internal.h"
#include "../../include/gtlattice.h"

namespace gt {
namespace {

constexpr size_t kChunk = (size_t)16 << 20;   // bytes per pipelined chunk

int type_size(int t) {
    switch (t) {
    case GT_TMP_F32: case GT_TMP_U32: case GT_TMP_I32: return 4;
    case GT_TMP_F64: case GT_TMP_U64: case GT_TMP_I64: return 8;
    case GT_TMP_U8: case GT_TMP_I8: return 1;
    case GT_TMP_U16: case GT_TMP_I16: return 2;
    default: return 0;
    }
}

template <typename T> __host__ __device__ inline float to_f32(T v) { return (float)v; }

__host__ __device__ inline int32_t trunc_i32(float x) {
    const float t = truncf(x);
    return (t >= -2147483648.0f && t < 2147483648.0f) ? (int32_t)t : INT32_MIN;
}
__host__ __device__ inline int64_t trunc_i64(float x) {
    const float t = truncf(x);
    return (t >= -9223372036854775808.0f && t < 9223372036854775808.0f) ? (int64_t)t : INT64_MIN;
}
template <typename T> __host__ __device__ inline T from_f32(float x);
template <> __host__ __device__ inline float from_f32<float>(float x) { return x; }
template <> __host__ __device__ inline double from_f32<double>(float x) { return (double)x; }
template <> __host__ __device__ inline uint8_t from_f32<uint8_t>(float x) { return (uint8_t)trunc_i32(x); }
template <> __host__ __device__ inline int8_t from_f32<int8_t>(float x) { return (int8_t)trunc_i32(x); }
template <> __host__ __device__ inline uint16_t from_f32<uint16_t>(float x) { return (uint16_t)trunc_i32(x); }
template <> __host__ __device__ inline int16_t from_f32<int16_t>(float x) { return (int16_t)trunc_i32(x); }
template <> __host__ __device__ inline int32_t from_f32<int32_t>(float x) { return trunc_i32(x); }
template <> __host__ __device__ inline uint32_t from_f32<uint32_t>(float x) { return (uint32_t)trunc_i64(x); }
template <> __host__ __device__ inline int64_t from_f32<int64_t>(float x) { return trunc_i64(x); }
template <> __host__ __device__ inline uint64_t from_f32<uint64_t>(float x) {
    const float t = truncf(x);
    if (t >= 9223372036854775808.0f && t < 18446744073709551616.0f)
        return (uint64_t)trunc_i64(t - 9223372036854775808.0f) ^ (1ull << 63);
    return (uint64_t)trunc_i64(x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_to_f32(const T *__restrict__ in, float *__restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = to_f32(in[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_from_f32(const float *__restrict__ in, T *__restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = from_f32<T>(in[i]);
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 4096); }

template <typename T>
void launch_to(const void *in, float *out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f32<T>, dim3(grid_for(n)), dim3(256), 0, s, (const T *)in, out, n);
}
template <typename T>
void launch_from(const float *in, void *out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_from_f32<T>, dim3(grid_for(n)), dim3(256), 0, s, in, (T *)out, n);
}

void convert_to_f32(int type, const void *in, float *out, int64_t n, hipStream_t s) {
    switch (type) {
    case GT_TMP_F64: launch_to<double>(in, out, n, s); break;
    case GT_TMP_U8: launch_to<uint8_t>(in, out, n, s); break;
    case GT_TMP_I8: launch_to<int8_t>(in, out, n, s); break;
    case GT_TMP_U16: launch_to<uint16_t>(in, out, n, s); break;
    case GT_TMP_I16: launch_to<int16_t>(in, out, n, s); break;
    case GT_TMP_U32: launch_to<uint32_t>(in, out, n, s); break;
    case GT_TMP_I32: launch_to<int32_t>(in, out, n, s); break;
    case GT_TMP_U64: launch_to<uint64_t>(in, out, n, s); break;
    case GT_TMP_I64: launch_to<int64_t>(in, out, n, s); break;
    default: break;
    }
}

void convert_from_f32(int type, const float *in, void *out, int64_t n, hipStream_t s) {
    switch (type) {
    case GT_TMP_F32: launch_from<float>(in, out, n, s); break;
    case GT_TMP_F64: launch_from<double>(in, out, n, s); break;
    case GT_TMP_U8: launch_from<uint8_t>(in, out, n, s); break;
    case GT_TMP_I8: launch_from<int8_t>(in, out, n, s); break;
    case GT_TMP_U16: launch_from<uint16_t>(in, out, n, s); break;
    case GT_TMP_I16: launch_from<int16_t>(in, out, n, s); break;
    case GT_TMP_U32: launch_from<uint32_t>(in, out, n, s); break;
    case GT_TMP_I32: launch_from<int32_t>(in, out, n, s); break;
    case GT_TMP_U64: launch_from<uint64_t>(in, out, n, s); break;
    case GT_TMP_I64: launch_from<int64_t>(in, out, n, s); break;
    default: break;
    }
}

template <typename T>
void host_to_f32(const void *in, float *out, size_t n) {
    const T *p = (const T *)in;
    for (size_t i = 0; i < n; i++) out[i] = to_f32(p[i]);
}
template <typename T>
void host_from_f32(const float *in, void *out, size_t n) {
    T *p = (T *)out;
    for (size_t i = 0; i < n; i++) p[i] = from_f32<T>(in[i]);
}

void host_convert_to_f32(int type, const void *in, float *out, size_t n) {
    switch (type) {
    case GT_TMP_F32: memcpy(out, in, n * 4); break;
    case GT_TMP_F64: host_to_f32<double>(in, out, n); break;
    case GT_TMP_U8: host_to_f32<uint8_t>(in, out, n); break;
    case GT_TMP_I8: host_to_f32<int8_t>(in, out, n); break;
    case GT_TMP_U16: host_to_f32<uint16_t>(in, out, n); break;
    case GT_TMP_I16: host_to_f32<int16_t>(in, out, n); break;
    case GT_TMP_U32: host_to_f32<uint32_t>(in, out, n); break;
    case GT_TMP_I32: host_to_f32<int32_t>(in, out, n); break;
    case GT_TMP_U64: host_to_f32<uint64_t>(in, out, n); break;
    case GT_TMP_I64: host_to_f32<int64_t>(in, out, n); break;
    default: break;
    }
}

void host_convert_from_f32(int type, const float *in, void *out, size_t n) {
    switch (type) {
    case GT_TMP_F32: memcpy(out, in, n * 4); break;
    case GT_TMP_F64: host_from_f32<double>(in, out, n); break;
    case GT_TMP_U8: host_from_f32<uint8_t>(in, out, n); break;
    case GT_TMP_I8: host_from_f32<int8_t>(in, out, n); break;
    case GT_TMP_U16: host_from_f32<uint16_t>(in, out, n); break;
    case GT_TMP_I16: host_from_f32<int16_t>(in, out, n); break;
    case GT_TMP_U32: host_from_f32<uint32_t>(in, out, n); break;
    case GT_TMP_I32: host_from_f32<int32_t>(in, out, n); break;
    case GT_TMP_U64: host_from_f32<uint64_t>(in, out, n); break;
    case GT_TMP_I64: host_from_f32<int64_t>(in, out, n); break;
    default: break;
    }
}

// An open .tmp file positioned at its first sample.  A type code outside
// 0..9 is an old-style file with a four-int header and float samples
// (src/FileTMP.cpp:210-214).
struct TmpFile {
    FILE *f = nullptr;
    int dims[4] = {0, 0, 0, 0};
    int type = 0;
    int64_t n = 0;   // samples
    ~TmpFile() {
        if (f) fclose(f);
    }
};

int open_tmp(const char *path, TmpFile &t) {
    if (!path) return set_error(GT_EINVAL, "null path");
    t.f = fopen(path, "rb");
    if (!t.f) return set_error(GT_EINVAL, (std::string("could not open file ") + path).c_str());
    int32_t h[5];
    if (fread(h, sizeof(int32_t), 5, t.f) != 5) return set_error(GT_EINVAL, "file ended before end of header");
    t.type = h[4];
    if (type_size(t.type) == 0) {   // old-style header
        t.type = GT_TMP_F32;
        if (fseek(t.f, 16, SEEK_SET) != 0) return set_error(GT_EINVAL, "seek failed");
    }
    for (int i = 0; i < 4; i++) {
        if (h[i] < 1) return set_error(GT_EINVAL, "bad .tmp size");
        t.dims[i] = h[i];
    }
    t.n = (int64_t)h[0] * h[1] * h[2] * h[3];
    return GT_OK;
}

void report_dims(const TmpFile &t, int *dims) {
    if (dims)
        for (int i = 0; i < 4; i++) dims[i] = t.dims[i];
}

int write_header(FILE *f, int W, int H, int F, int C, int type) {
    const int32_t h[5] = {W, H, F, C, type};
    return fwrite(h, sizeof(int32_t), 5, f) == 5 ? GT_OK : set_error(GT_EDEVICE, "write failed");
}

// Two pinned host buffers (and device staging) for the pipelined copies.
// The destructor first waits for the stream, so no return path frees a
// buffer that a copy in flight still uses.
struct Pinned {
    void *p[2] = {nullptr, nullptr};
    void *d[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;
    ~Pinned() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (int i = 0; i < 2; i++) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (p[i]) (void)hipHostFree(p[i]);
            if (d[i]) (void)hipFree(d[i]);
        }
    }
    int init(bool device_staging) {
        for (int i = 0; i < 2; i++) {
            if (alloc_should_fail() || hipHostMalloc(&p[i], kChunk, hipHostMallocDefault) != hipSuccess ||
                (device_staging && hipMalloc(&d[i], kChunk) != hipSuccess) ||
                hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                return set_error(GT_ENOMEM, "pinned staging allocation failed");
            }
        }
        return GT_OK;
    }
};

} // namespace
} // namespace gt

using namespace gt;

extern "C" {

int gt_tmp_info(const char *path, int *dims, int *type) {
    TmpFile t;
    const int rc = open_tmp(path, t);
    if (rc != GT_OK) return rc;
    report_dims(t, dims);
    if (type) *type = t.type;
    return GT_OK;
}

int gt_tmp_read(const char *path, float *out, size_t capacity, int *dims) {
    TmpFile t;
    int rc = open_tmp(path, t);
    if (rc != GT_OK) return rc;
    report_dims(t, dims);
    if (!out) return GT_OK;
    if ((size_t)t.n > capacity) return set_error(GT_EINVAL, "output buffer too small for the image");
    const int ts = type_size(t.type);
    std::vector<char> buf(std::min<size_t>(kChunk, (size_t)t.n * ts));
    const size_t per = buf.size() / ts;
    for (int64_t done = 0; done < t.n;) {
        const size_t k = (size_t)std::min<int64_t>(per, t.n - done);
        if (fread(buf.data(), ts, k, t.f) != k) return set_error(GT_EINVAL, "unexpected end of file");
        host_convert_to_f32(t.type, buf.data(), out + done, k);
        done += k;
    }
    return GT_OK;
}

int gt_tmp_write(const char *path, const float *im, int W, int H, int F, int C, int type) {
    if (!path || !im) return set_error(GT_EINVAL, "null pointer");
    if (W < 1 || H < 1 || F < 1 || C < 1) return set_error(GT_EINVAL, "bad image size");
    const int ts = type_size(type);
    if (!ts) return set_error(GT_EINVAL, "unknown .tmp type code");
    FILE *f = fopen(path, "wb");
    if (!f) return set_error(GT_EINVAL, (std::string("could not write output file ") + path).c_str());
    int rc = write_header(f, W, H, F, C, type);
    const int64_t n = (int64_t)W * H * F * C;
    std::vector<char> buf(std::min<size_t>(kChunk, (size_t)n * ts));
    const size_t per = buf.size() / ts;
    for (int64_t done = 0; rc == GT_OK && done < n;) {
        const size_t k = (size_t)std::min<int64_t>(per, n - done);
        host_convert_from_f32(type, im + done, buf.data(), k);
        if (fwrite(buf.data(), ts, k, f) != k) rc = set_error(GT_EDEVICE, "write failed");
        done += k;
    }
    if (fclose(f) != 0 && rc == GT_OK) rc = set_error(GT_EDEVICE, "close failed");
    return rc;
}

int gt_tmp_load(gt_ctx *ctx, const char *path, float *dev_out, size_t capacity, int *dims) {
    if (!ctx) return set_error(GT_EINVAL, "null context");
    TmpFile t;
    int rc = open_tmp(path, t);
    if (rc != GT_OK) return rc;
    report_dims(t, dims);
    if (!dev_out) return GT_OK;
    if ((size_t)t.n > capacity) return set_error(GT_EINVAL, "output buffer too small for the image");
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return set_error(GT_EDEVICE, "device unavailable");
    hipStream_t s = ctx_stream(ctx);
    const int ts = type_size(t.type);
    const bool raw = t.type == GT_TMP_F32;
    Pinned pin;
    pin.stream = s;
    if ((rc = pin.init(!raw)) != GT_OK) return rc;
    const size_t per = kChunk / ts;
    int64_t done = 0;
    for (int i = 0; done < t.n; i ^= 1) {
        const size_t k = (size_t)std::min<int64_t>(per, t.n - done);
        if (hipEventSynchronize(pin.ev[i]) != hipSuccess) return set_error(GT_EDEVICE, "staging wait failed");
        if (fread(pin.p[i], ts, k, t.f) != k) {
            (void)hipStreamSynchronize(s);
            return set_error(GT_EINVAL, "unexpected end of file");
        }
        void *dst = raw ? (void *)(dev_out + done) : pin.d[i];
        if (hipMemcpyAsync(dst, pin.p[i], k * ts, hipMemcpyHostToDevice, s) != hipSuccess)
            return set_error(GT_EDEVICE, "copy to device failed");
        if (!raw) convert_to_f32(t.type, pin.d[i], dev_out + done, (int64_t)k, s);
        if (hipEventRecord(pin.ev[i], s) != hipSuccess) return set_error(GT_EDEVICE, "event record failed");
        done += k;
    }
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
        return set_error(GT_EDEVICE, "load failed on the device");
    return GT_OK;
}

int gt_tmp_save(gt_ctx *ctx, const char *path, const float *dev_im, int W, int H, int F, int C, int type) {
    if (!ctx || !path || !dev_im) return set_error(GT_EINVAL, "null pointer");
    if (W < 1 || H < 1 || F < 1 || C < 1) return set_error(GT_EINVAL, "bad image size");
    const int ts = type_size(type);
    if (!ts) return set_error(GT_EINVAL, "unknown .tmp type code");
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return set_error(GT_EDEVICE, "device unavailable");
    hipStream_t s = ctx_stream(ctx);
    Pinned pin;
    pin.stream = s;
    int rc = pin.init(true);
    if (rc != GT_OK) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return set_error(GT_EINVAL, (std::string("could not write output file ") + path).c_str());
    rc = write_header(f, W, H, F, C, type);
    const int64_t n = (int64_t)W * H * F * C;
    const size_t per = kChunk / std::max(ts, 4);
    // chunk i is converted and copied into pinned buffer i&1 while chunk i-1 is written to disk
    int64_t pending = -1;
    size_t pending_k = 0;
    int pi = 0;
    for (int64_t done = 0, i = 0; rc == GT_OK && (done < n || pending >= 0); i++) {
        int64_t issued = -1;
        size_t k = 0;
        const int b = (int)(i & 1);
        if (done < n) {
            k = (size_t)std::min<int64_t>(per, n - done);
            convert_from_f32(type, dev_im + done, pin.d[b], (int64_t)k, s);
            if (hipMemcpyAsync(pin.p[b], pin.d[b], k * ts, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipEventRecord(pin.ev[b], s) != hipSuccess)
                rc = set_error(GT_EDEVICE, "copy from device failed");
            issued = done;
            done += k;
        }
        if (rc == GT_OK && pending >= 0) {
            if (hipEventSynchronize(pin.ev[pi]) != hipSuccess) rc = set_error(GT_EDEVICE, "staging wait failed");
            else if (fwrite(pin.p[pi], ts, pending_k, f) != pending_k) rc = set_error(GT_EDEVICE, "write failed");
        }
        pending = issued;
        pending_k = k;
        pi = b;
    }
    (void)hipStreamSynchronize(s);
    if (fclose(f) != 0 && rc == GT_OK) rc = set_error(GT_EDEVICE, "close failed");
    return rc;
}

} // extern "C"
