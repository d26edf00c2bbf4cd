// C ABI of add_noise (crates/lucem-ops/src/intensity/noise.rs:35-143): the six noise models on
// the device, with the reference's seed folding and parameter handling.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

extern "C" lucem_status lucem_add_noise_f32(lucem_client* client, const float* input, const uint64_t* dims,
                                            uint32_t ndim, float* output, const lucem_noise_model* model,
                                            uint64_t seed, int32_t clip) {
    if (!model) return fail(LUCEM_INVALID_HANDLE, "noise model is NULL");
    if (ndim && !dims) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < ndim; i++) {
        if (dims[i] && n > UINT64_MAX / dims[i]) overflow = true;
        n *= dims[i];
    }
    if (n == 0) return fail(LUCEM_EMPTY_INPUT, "empty input to add_noise");
    if (model->kind > LUCEM_NOISE_POISSON)
        return fail(LUCEM_INVALID_PARAMETER, "invalid parameter `model`: unknown noise model " +
                                                 std::to_string(model->kind));
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(dims, ndim) + " exceed the device element limit of 4294967295");
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!input || !output) return fail(LUCEM_INVALID_HANDLE, "input or output pointer is NULL");
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // noise.rs:46: the 64-bit seed folded to 32 bits
    const uint32_t seed32 = (uint32_t)seed ^ ((uint32_t)(seed >> 32) * 0x9E3779B9u);
    float a = (float)model->p0, b = (float)model->p1;
    switch (model->kind) {
        case LUCEM_NOISE_SALT: b = 1.0f; break;    // every hit is salt (noise.rs:80-93)
        case LUCEM_NOISE_PEPPER: b = 0.0f; break;  // every hit is pepper (:94-107)
        case LUCEM_NOISE_POISSON: a = (float)(model->p0 > 1.0 ? model->p0 : 1.0); break;  // scale.max(1) (:134)
        default: break;
    }
    const int err = launch_add_noise(input, output, (int)model->kind, seed32, a, b, clip ? 1 : 0, (uint32_t)n,
                                     client->stream);
    return err == hipSuccess ? LUCEM_OK : hip_fail(err, "add_noise");
}
