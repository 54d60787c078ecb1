// Runtime entry points of libnca_hip.so: errors, contexts, point sets.
#include <algorithm>
#include <memory>

#include "nca_common.hpp"

namespace nca {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace nca

using namespace nca;

extern "C" {

const char* nca_last_error(void) { return g_last_error.c_str(); }

int32_t nca_abi_version(void) { return 1; }

int32_t nca_ctx_create(int32_t device, nca_ctx** out) {
    return guarded([&] {
        NCA_REQUIRE(out, NCA_ERR_ARGUMENT, "nca_ctx_create: null output");
        int count = 0;
        NCA_HIP(hipGetDeviceCount(&count));
        NCA_REQUIRE(device >= 0 && device < count, NCA_ERR_ARGUMENT,
                    "nca_ctx_create: device " + std::to_string(device) + " of " + std::to_string(count));
        auto ctx = std::make_unique<nca_ctx>();
        ctx->device = device;
        NCA_HIP(hipSetDevice(device));
        NCA_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        *out = ctx.release();
    });
}

int32_t nca_ctx_destroy(nca_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        set_device(ctx);
        if (ctx->stream) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipStreamDestroy(ctx->stream);
        }
        delete ctx;
    });
}

int32_t nca_ctx_synchronize(nca_ctx* ctx) {
    return guarded([&] {
        NCA_REQUIRE(ctx, NCA_ERR_ARGUMENT, "nca_ctx_synchronize: null context");
        set_device(ctx);
        NCA_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int32_t nca_ctx_stream(nca_ctx* ctx, void** stream) {
    return guarded([&] {
        NCA_REQUIRE(ctx && stream, NCA_ERR_ARGUMENT, "nca_ctx_stream: invalid argument");
        *stream = (void*)ctx->stream;
    });
}

int32_t nca_ctx_device_name(nca_ctx* ctx, char* buf, int64_t len) {
    return guarded([&] {
        NCA_REQUIRE(ctx && buf && len > 0, NCA_ERR_ARGUMENT, "nca_ctx_device_name: invalid argument");
        hipDeviceProp_t prop;
        NCA_HIP(hipGetDeviceProperties(&prop, ctx->device));
        std::string name = prop.gcnArchName;
        const size_t n = std::min<size_t>(name.size(), (size_t)len - 1);
        std::memcpy(buf, name.data(), n);
        buf[n] = 0;
    });
}

int32_t nca_points_create(nca_ctx* ctx, const double* xyz, int64_t n, nca_points** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && out && n >= 0 && (n == 0 || xyz), NCA_ERR_ARGUMENT, "nca_points_create: invalid argument");
        set_device(ctx);
        auto p = std::make_unique<nca_points>();
        p->ctx = ctx;
        p->n = n;
        p->host.assign(xyz, xyz + 3 * n);
        p->xyz.upload(p->host, ctx->stream);
        NCA_HIP(hipStreamSynchronize(ctx->stream));
        *out = p.release();
    });
}

int32_t nca_points_destroy(nca_points* pts) {
    return guarded([&] {
        if (!pts) return;
        if (pts->ctx) set_device(pts->ctx);
        delete pts;
    });
}

}  // extern "C"
