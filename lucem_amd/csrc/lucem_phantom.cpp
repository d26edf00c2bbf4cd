// C ABI of the device phantoms: the synthetic workloads' ground truth (background plus filled
// ellipses or spheres) rendered on the device, so that a workload can be generated without
// moving images from the host. No reference counterpart.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lucem_hip.h"
#include "lucem_internal_host.hpp"
#include "rl_internal.hpp"

using namespace lucem;

namespace {

// Checks shared by both renderers: rank, empty output, element limit, the launch grid's limits,
// then the handles.
lucem_status validate_phantom(lucem_client* client, const float* out, const uint64_t* dims, uint32_t ndim,
                              uint32_t want_ndim, const void* objects, uint64_t count, const uint64_t* max_dims,
                              bool* empty) {
    *empty = false;
    if (ndim && !dims) return fail(LUCEM_INVALID_HANDLE, "dims pointer is NULL");
    if (ndim != want_ndim)
        return fail(LUCEM_INVALID_DIMENSION, std::string(want_ndim == 2 ? "render_ellipses" : "render_spheres") +
                                                 " expects " + std::to_string(want_ndim) + "-D array, got " +
                                                 std::to_string(ndim) + "-D shape " + dims_str(dims, ndim));
    uint64_t n = 1;
    bool overflow = false;
    for (uint32_t i = 0; i < ndim; i++) {
        if (dims[i] && n > UINT64_MAX / dims[i]) overflow = true;
        n *= dims[i];
    }
    if (n == 0) {
        *empty = true;
        return LUCEM_OK;
    }
    if (overflow || n > 0xFFFFFFFFull)
        return fail(LUCEM_TOO_MANY_ELEMENTS,
                    "dims " + dims_str(dims, ndim) + " exceed the device element limit of 4294967295");
    for (uint32_t i = 0; i < ndim; i++)
        if (max_dims[i] && dims[i] > max_dims[i])
            return fail(LUCEM_UNSUPPORTED_SHAPE, "dims " + dims_str(dims, ndim) + ": axis " + std::to_string(i) +
                                                     " longer than " + std::to_string(max_dims[i]));
    if (!client) return fail(LUCEM_INVALID_HANDLE, "client is NULL");
    if (!out) return fail(LUCEM_INVALID_HANDLE, "output pointer is NULL");
    if (count && !objects) return fail(LUCEM_INVALID_HANDLE, "objects pointer is NULL");
    return LUCEM_OK;
}

}  // namespace

extern "C" lucem_status lucem_render_ellipses_f32(lucem_client* client, float* out, const uint64_t* dims,
                                                  uint32_t ndim, const lucem_ellipse* objects, uint64_t count,
                                                  double background) {
    bool empty = false;
    const uint64_t max_dims[2] = {65535ull * 16, 0};  // grid y: 16 rows per workgroup
    lucem_status st = validate_phantom(client, out, dims, ndim, 2, objects, count, max_dims, &empty);
    if (st != LUCEM_OK || empty) return st;
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int err = launch_render_ellipses(out, (int)dims[0], (int)dims[1], objects, (int64_t)count, background,
                                           client->stream);
    return err == hipSuccess ? LUCEM_OK : hip_fail(err, "render ellipses");
}

extern "C" lucem_status lucem_render_spheres_f32(lucem_client* client, float* out, const uint64_t* dims, uint32_t ndim,
                                                 const lucem_sphere* objects, uint64_t count, double background) {
    bool empty = false;
    const uint64_t max_dims[3] = {65535ull * 4, 65535ull * 8, 0};  // grid z: 4 planes, grid y: 8 rows
    lucem_status st = validate_phantom(client, out, dims, ndim, 3, objects, count, max_dims, &empty);
    if (st != LUCEM_OK || empty) return st;
    hipError_t e = hipSetDevice(client->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int err = launch_render_spheres(out, (int)dims[0], (int)dims[1], (int)dims[2], objects, (int64_t)count,
                                          background, client->stream);
    return err == hipSuccess ? LUCEM_OK : hip_fail(err, "render spheres");
}
