// The kernels' boundary index map (lucem::pad_index in lucem_amd/csrc/fft_lds.hpp), callable from
// the host through ctypes so that tests/test_kernel_math.py can compare it with the oracle's
// restatement of crates/lucem-ops/src/core/kernels/pad.rs:11-34 position by position.
#include "../../lucem_amd/csrc/fft_lds.hpp"

extern "C" void lucem_probe_pad_index(const int* p, int count, int len, int mode, int* out) {
    for (int i = 0; i < count; i++) out[i] = lucem::pad_index(p[i], len, mode);
}
