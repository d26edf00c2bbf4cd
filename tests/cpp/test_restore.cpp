// C++ caller test: drives the library exactly as a host-language binding would, through the
// C++ mirror of lucem-ops' interface (include/lucem/restore.hpp). Mirrors the reference test
// crates/lucem-ops/src/restore/deconvolve.rs:517-566 (identity KAT within 1e-4, box-PSF peak
// gain > 0.15) and the error order of deconvolve.rs:53-71, plus the Wiener, noise and correlation
// wrappers. Prints "ok" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/lucem/restore.hpp"

using namespace lucem;

static int failures = 0;
#define CHECK(cond, ...)                              \
    do {                                              \
        if (!(cond)) {                                \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);        \
            std::fprintf(stderr, "\n");               \
            failures++;                               \
        }                                             \
    } while (0)

int main() {
    Client client(0);
    // identity: 8x8 ramp, delta PSF, 15 iterations, no clip
    std::vector<float> img(64);
    for (int i = 0; i < 64; i++) img[i] = (float(i) / 64.0f) * 0.8f + 0.1f;
    std::vector<float> delta(9, 0.0f);
    delta[4] = 1.0f;
    auto in = DeviceArray<float>::from_host(client, {8, 8}, img).unwrap();
    auto psf = DeviceArray<float>::from_host(client, {3, 3}, delta).unwrap();
    RichardsonLucy s;
    s.num_iter = 15;
    s.clip = false;
    auto out = richardson_lucy(client, in, psf, s).unwrap().to_host(client);
    double scale = 0;
    for (float v : img) scale = std::fmax(scale, std::fabs(v));
    for (int i = 0; i < 64; i++) {
        const double tol = 1e-4 * std::fabs(img[i]) + 1e-4 * scale;
        CHECK(std::fabs(out[i] - img[i]) <= tol, "identity index %d: %f vs %f", i, out[i], img[i]);
    }
    // sharpening: spot 0.9 on 0.1, box-blurred (the blur is RL's own forward operator with a
    // symmetric PSF), then 25 iterations with clip on
    std::vector<float> spot(64, 0.1f), blurred(64, 0.0f);
    spot[36] = 0.9f;
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) {
            float acc = 0;
            for (int ky = 0; ky < 3; ky++)
                for (int kx = 0; kx < 3; kx++) {
                    int py = y - 1 + ky, px = x - 1 + kx;
                    py = py < 0 ? -py - 1 : (py > 7 ? 15 - py : py);
                    px = px < 0 ? -px - 1 : (px > 7 ? 15 - px : px);
                    acc += (1.0f / 9.0f) * spot[py * 8 + px];
                }
            blurred[y * 8 + x] = acc;
        }
    auto bl = DeviceArray<float>::from_host(client, {8, 8}, blurred).unwrap();
    auto box = DeviceArray<float>::from_host(client, {3, 3}, std::vector<float>(9, 1.0f / 9.0f)).unwrap();
    RichardsonLucy s2;
    s2.num_iter = 25;
    s2.clip = true;
    auto restored = richardson_lucy(client, bl, box, s2).unwrap().to_host(client);
    CHECK(restored[36] > blurred[36] + 0.15f, "peak %f vs blurred %f", restored[36], blurred[36]);
    // the f64 instance: the identity KAT in FP64 holds to rounding
    std::vector<double> img64(img.begin(), img.end()), delta64(delta.begin(), delta.end());
    auto in64 = DeviceArray<double>::from_host(client, {8, 8}, img64).unwrap();
    auto psf64 = DeviceArray<double>::from_host(client, {3, 3}, delta64).unwrap();
    auto out64 = richardson_lucy(client, in64, psf64, s).unwrap().to_host(client);
    for (int i = 0; i < 64; i++)
        CHECK(std::fabs(out64[i] - img64[i]) <= 1e-12, "f64 identity index %d: %.17g vs %.17g", i, out64[i], img64[i]);
    // Wiener: the identity KAT of deconvolve.rs:542-550 (delta PSF, balance 1e-7, within 1e-3)
    WienerDeconvolve wsettings;
    wsettings.balance = 1e-7;
    wsettings.clip = false;
    auto wout = wiener_deconvolve(client, in, psf, wsettings).unwrap().to_host(client);
    for (int i = 0; i < 64; i++) CHECK(std::fabs(wout[i] - img[i]) <= 1e-3, "wiener identity index %d", i);
    // unsupervised Wiener on the blurred spot (deconvolve.rs:569-578): a chain of 12, an 8x8 image
    UnsupervisedWiener us;
    us.num_iter = 12;
    us.burnin = 4;
    us.clip = true;
    us.seed = 7;
    auto uw = unsupervised_wiener(client, bl, box, us).unwrap();
    CHECK(uw.second.size() == 12, "chain length %zu", uw.second.size());
    CHECK(uw.first.dims().size() == 2 && uw.first.dims()[0] == 8 && uw.first.dims()[1] == 8, "uw dims");
    // add_noise: clipped Gaussian noise stays in [0, 1] (intensity/mod.rs:52-63)
    auto noisy = add_noise(client, in, NoiseModel::gaussian(0.0, 0.05), 42, true).unwrap().to_host(client);
    for (float v : noisy) CHECK(v >= 0.0f && v <= 1.0f, "noise out of range %f", v);
    // correlate_2d: a delta window at (2, 0) of a 3x3 window under Wrap reads in[y + 1, x - 1]
    // (filter/convolve.rs:126-166, left = k / 2 + origin); f32 through the FFT passes, f64 direct
    {
        std::vector<float> dw(9, 0.0f);
        dw[6] = 1.0f;
        auto win = DeviceArray<float>::from_host(client, {3, 3}, dw).unwrap();
        Correlate2d cs;
        cs.mode = PadMode::Wrap;
        auto c32 = correlate_2d(client, in, win, cs).unwrap().to_host(client);
        std::vector<double> img64(img.begin(), img.end()), dw64(dw.begin(), dw.end());
        auto in64 = DeviceArray<double>::from_host(client, {8, 8}, img64).unwrap();
        auto win64 = DeviceArray<double>::from_host(client, {3, 3}, dw64).unwrap();
        auto c64 = correlate_2d(client, in64, win64, cs).unwrap().to_host(client);
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) {
                const double want = img[((y + 1) % 8) * 8 + (x + 7) % 8];
                CHECK(std::fabs(c32[y * 8 + x] - want) < 1e-6, "correlate f32 (%d, %d)", y, x);
                CHECK(c64[y * 8 + x] == want, "correlate f64 (%d, %d)", y, x);
            }
        cs.origin = {2, 0};
        CHECK(correlate_2d(client, in, win, cs).status() == LUCEM_INVALID_ORIGIN, "origin check");
    }
    // phase_correlation: an image against itself registers at zero shift (register/phase.rs:52-250)
    {
        auto pc = phase_correlation(client, in, in, PhaseCorrelation{4, true}).unwrap();
        CHECK(pc.shift.size() == 2 && pc.shift[0] == 0.0 && pc.shift[1] == 0.0, "phase correlation shift");
        CHECK(std::fabs(pc.phasediff) < 1e-5, "phase correlation phase %f", pc.phasediff);
    }
    // butterworth: the low and high passes add up to the input (filter/spectral.rs:166-212)
    {
        auto lo = butterworth(client, in, 0.2, false, 2.0).unwrap().to_host(client);
        auto hi = butterworth(client, in, 0.2, true, 2.0).unwrap().to_host(client);
        for (int i = 0; i < 64; i++) CHECK(std::fabs(lo[i] + hi[i] - img[i]) < 1e-5, "butterworth sum %d", i);
        auto gb = gabor_2d(client, in, 0.1, 0.0, 1.0, 1.0, 3.0, 0.0, PadMode::Reflect, 0.0);
        CHECK(gb.is_ok(), "gabor_2d");
    }
    // error order
    Workspace ws(client);
    auto out86 = DeviceArray<float>::empty(client, {8, 6}).unwrap();
    auto r = richardson_lucy_into(client, in, psf, out86, s, ws);
    CHECK(r.status() == LUCEM_SHAPE_MISMATCH, "shape mismatch status %d", (int)r.status());
    r = richardson_lucy_into(client, in, psf, in, s, ws);
    CHECK(r.status() == LUCEM_ALIASED_BUFFERS, "aliased status %d", (int)r.status());
    RichardsonLucy s0 = s;
    s0.num_iter = 0;
    auto out88 = DeviceArray<float>::empty(client, {8, 8}).unwrap();
    r = richardson_lucy_into(client, in, psf, out88, s0, ws);
    CHECK(r.status() == LUCEM_INVALID_PARAMETER, "num_iter status %d", (int)r.status());
    auto pm = parse_pad_mode("grid-wrap");
    CHECK(pm.is_ok() && pm.unwrap() == PadMode::Wrap, "parse grid-wrap");
    CHECK(!parse_pad_mode("nope").is_ok(), "parse rejects");
    client.sync();
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
