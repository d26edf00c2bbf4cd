// Launch-shape invariants of the planning functions, on the host (no device is touched): for
// every 7-smooth FFT length up to 16384, the row and line shapes the heuristics choose fit the
// thread tier they name (lines x L <= threads x 16), the radix plans multiply out to L, and the
// general FFT's axis shapes (Stockham or Bluestein) fit theirs. Prints "ok" and exits 0.
#include <cstdio>

#include "../../lucem_amd/csrc/lucem_internal_host.hpp"

using namespace lucem;

static int failures = 0;

static void check(bool ok, const char* what, int L, int a, int b) {
    if (!ok && failures++ < 20) std::fprintf(stderr, "FAIL %s: L=%d (%d, %d)\n", what, L, a, b);
}

static bool smooth(int v) {
    for (int p : {2, 3, 5, 7})
        while (v % p == 0) v /= p;
    return v == 1;
}

int main() {
    int lengths = 0;
    for (int L = 1; L <= LUCEM_MAX_FFT_LEN; L++) {
        if (!smooth(L)) continue;
        lengths++;
        for (int variant = 0; variant < 3; variant++) {
            const FftPlan f = make_plan(L, variant);
            int prod = 1;
            for (int q = 0; q < f.npass; q++) {
                const int code = plan_code(f, q);
                prod *= (code & 15) * (code >> 4);
            }
            check(prod == L, "plan product", L, variant, prod);
        }
        for (int num_cus : {0, 256}) {
            for (int rows : {1, 2, 7, 512, 100000}) {
                const ShapePolicy sp{num_cus, 0, 0, nullptr};
                int tier = -1, var = 0;
                const int P = choose_P(L, rows, sp, &tier, &var);
                check(P >= 1 && tier >= 0 && tier <= 2 && (long)P * L <= (long)tier_threads(tier) * kElemsPerThread,
                      "row shape", L, P, tier);
            }
            for (int kind : {SHAPE_COL_FWD, SHAPE_COL_FUSED, SHAPE_COL_INV})
                for (int Lc : {1, 9, 271, 8193})
                    for (long outers : {1L, 128L, 4096L}) {
                        const ShapePolicy sp{num_cus, 0, 0, nullptr};
                        int wide = -1, var = 0;
                        const int B = choose_B(L, Lc, outers, kind, sp, &wide, &var);
                        check(B >= 1 && wide >= 0 && wide <= 2 &&
                                  (long)B * L <= (long)tier_threads(wide) * kElemsPerThread,
                              "line shape", L, B, wide);
                    }
        }
    }
    std::printf("%s\n", failures ? "failed" : "ok");
    std::fprintf(stderr, "%d lengths checked\n", lengths);
    return failures ? 1 : 0;
}
