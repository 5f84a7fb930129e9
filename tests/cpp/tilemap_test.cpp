// CPU test of the workgroup-to-tile maps of the sweep launches
// (physicell-x_amd/csrc/biofvm_tilemap.h): for every launch size G up to
// 2100 and the sizes of the 400^3 and 800^3 column sweeps, and R from 1 to
// 70, xcd_runs is a bijection on [0, G); with R = 1 it equals xcd_block;
// the workgroups of one XCD (ids b = x mod 8) cover one contiguous stretch
// of tiles, as XCD x's share of xcd_block; and within that share, the
// workgroups with the same i mod R (i = b / 8) take consecutive tiles in
// order.  Exit status 0 on success.
#include "biofvm_tilemap.h"

#include <cstdio>
#include <vector>

static_assert(bfa::xcd_block(0, 16) == 0 && bfa::xcd_block(1, 16) == 2 && bfa::xcd_block(8, 16) == 1,
              "two tiles per XCD, in order");
static_assert(bfa::xcd_runs(8, 32, 2) == 2 && bfa::xcd_runs(16, 32, 2) == 1,
              "four tiles per XCD in two runs: workgroup i = 1 starts the second run");

static int check(int G, int R) {
    std::vector<int> seen(G, 0);
    for (int b = 0; b < G; ++b) {
        const int t = bfa::xcd_runs(b, G, R);
        if (t < 0 || t >= G) return 1;
        if (seen[t]++) return 2;
        if (R == 1 && t != bfa::xcd_block(b, G)) return 3;
    }
    for (int x = 0; x < 8 && x < G; ++x) {
        // XCD x: the same share as xcd_block, split into runs in order
        int lo = G, hi = -1, lo1 = G, hi1 = -1;
        for (int b = x; b < G; b += 8) {
            const int t = bfa::xcd_runs(b, G, R), t1 = bfa::xcd_block(b, G);
            lo = t < lo ? t : lo;
            hi = t > hi ? t : hi;
            lo1 = t1 < lo1 ? t1 : lo1;
            hi1 = t1 > hi1 ? t1 : hi1;
            const int i = b >> 3;
            if (i >= R) {
                const int prev = bfa::xcd_runs(b - 8 * R, G, R);  // same sub-run, one position back
                if (t != prev + 1) return 4;
            }
        }
        const int count = (G - x + 7) / 8;
        if (lo != lo1 || hi != hi1 || hi - lo + 1 != count) return 5;
    }
    return 0;
}

int main() {
    int failures = 0, cases = 0;
    std::vector<int> sizes;
    for (int G = 1; G <= 2100; ++G) sizes.push_back(G);
    for (int G : {20000, 40000, 10000, 80000, 12500, 25000}) sizes.push_back(G);
    for (int G : sizes) {
        for (int R = 1; R <= 70; ++R) {
            if (G > 2100 && R > 16 && R != 25 && R != 64) continue;
            const int rc = check(G, R);
            ++cases;
            if (rc) {
                if (++failures < 10) std::printf("G=%d R=%d: check %d failed\n", G, R, rc);
            }
        }
    }
    std::printf("%d cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
