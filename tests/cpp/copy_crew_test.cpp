// CPU test of the host-thread crew of the pinned transfer pipeline
// (physicell-x_amd/csrc/biofvm_transfer.inc): for crews of 1 to 16 threads
// and sizes around the 4 KiB slice alignment, one crew reused for every
// copy reproduces the source byte for byte and writes nothing outside the
// destination range.  No GPU call is made.  Exit status 0 on success.
#include "biofvm_transfer.inc"

#include <cstdio>
#include <random>

int main() {
    std::mt19937_64 rng(20260);
    const size_t sizes[] = {0, 1, 7, 4095, 4096, 4097, 3 * 65536 + 17, (size_t)5 << 20,
                            ((size_t)9 << 20) + 3};
    const size_t guard = 64;
    int failures = 0, copies = 0;
    for (int threads : {1, 2, 3, 8, 16}) {
        CopyCrew crew(threads);
        for (int rep = 0; rep < 2; ++rep) {
            for (size_t bytes : sizes) {
                std::vector<unsigned char> src(bytes + 2 * guard), dst(bytes + 2 * guard, 0xA5);
                for (auto& c : src) c = (unsigned char)rng();
                crew.copy(dst.data() + guard, src.data() + guard, bytes);
                ++copies;
                bool ok = std::memcmp(dst.data() + guard, src.data() + guard, bytes) == 0;
                for (size_t i = 0; i < guard; ++i)
                    ok = ok && dst[i] == 0xA5 && dst[guard + bytes + i] == 0xA5;
                if (!ok) {
                    std::printf("FAIL: %d threads, %zu bytes, repeat %d\n", threads, bytes, rep);
                    ++failures;
                }
            }
        }
    }
    std::printf("copy crew: %d copies, %d failures\n", copies, failures);
    return failures != 0;
}
