// CPU test of the host-thread crew of the pinned transfer pipeline
// (physicell-x_amd/csrc/biofvm_transfer.inc): for crews of 1 to 16 threads
// and sizes around the 4 KiB slice alignment, one crew reused for every
// copy reproduces the source byte for byte and writes nothing outside the
// destination range; element jobs visit every element once; the substrate
// runs of the converting transfers give e % ns / e / N.  No GPU call is
// made.  Exit status 0 on success.
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
    // element jobs: every element of [0, count) is visited exactly once
    int jobs = 0;
    for (int threads : {1, 2, 3, 8, 16}) {
        CopyCrew crew(threads);
        for (size_t count : {(size_t)0, (size_t)1, (size_t)1023, (size_t)1024, (size_t)1025, (size_t)5000,
                             ((size_t)3 << 20) + 5}) {
            std::vector<int> hits(count, 0);
            crew.each(count, [&](size_t a, size_t b) {
                for (size_t q = a; q < b; ++q) ++hits[q];
            });
            ++jobs;
            for (size_t q = 0; q < count; ++q)
                if (hits[q] != 1) {
                    std::printf("FAIL: job of %zu elements, %d threads: element %zu visited %d times\n", count,
                                threads, q, hits[q]);
                    ++failures;
                    break;
                }
        }
    }
    // runs of the converting transfers, from any start: contiguous, and the
    // substrate of each element is e % ns (reference layout) or e / N
    // (device layout)
    int walks = 0;
    for (int ns : {1, 2, 3, 7})
        for (long long N : {1LL, 5LL, 13LL})
            for (bool ref : {true, false}) {
                const size_t total = (size_t)N * ns;
                for (size_t e0 = 0; e0 < total; ++e0) {
                    size_t next = 0;
                    bool ok = true;
                    each_run(e0, total - e0, ns, N, ref, [&](size_t q, size_t m, int d) {
                        ok = ok && q == next && m > 0;
                        for (size_t k = 0; k < m; ++k) {
                            const size_t e = e0 + q + k;
                            const int want = ref ? (int)(e % ns) : (int)(e / N);
                            ok = ok && (ref ? d + (int)k : d) == want;
                        }
                        next = q + m;
                    });
                    if (!ok || next != total - e0) {
                        std::printf("FAIL: runs ns=%d N=%lld ref=%d from %zu\n", ns, N, (int)ref, e0);
                        ++failures;
                    }
                    ++walks;
                }
            }
    // the conversions of the FP32 transfers against their element-wise
    // definitions, bit for bit: u = fma(v, s[d], b[d]) and
    // v = (float)((u - b[d]) * rs[d]), for slices at any offset, length and
    // destination alignment, in both layouts
    int convs = 0;
    std::uniform_real_distribution<double> ud(-50.0, 50.0);
    for (int ns : {1, 2, 3, 4, 7})
        for (long long N : {5LL, 13LL, 1000LL})
            for (bool ref : {true, false}) {
                const size_t total = (size_t)N * ns;
                std::vector<double> s(ns), b(ns), rs(ns);
                for (int d = 0; d < ns; ++d) {
                    s[d] = std::ldexp(1.0, (int)(rng() % 9) - 4) * (d % 2 ? 1.0 : 1.5);
                    b[d] = ud(rng);
                    rs[d] = 1.0 / s[d];
                }
                std::vector<float> v(total);
                std::vector<double> u(total);
                for (size_t q = 0; q < total; ++q) {
                    v[q] = (float)ud(rng);
                    u[q] = ud(rng);
                }
                for (size_t e0 : {(size_t)0, (size_t)1, (size_t)3, total / 3})
                    for (size_t pad : {(size_t)0, (size_t)1, (size_t)2, (size_t)3}) {
                        if (e0 > total) continue;
                        const size_t n = total - e0;
                        std::vector<double> out(n + pad + 8, -1.0);
                        expand_slice(out.data() + pad, v.data() + e0, e0, n, s.data(), b.data(), ns, N, ref);
                        std::vector<float> back(n + 8, -1.0f);
                        compress_slice(back.data(), u.data() + e0, e0, n, b.data(), rs.data(), ns, N, ref);
                        bool ok = out[n + pad] == -1.0;
                        for (size_t q = 0; q < n && ok; ++q) {
                            const size_t e = e0 + q;
                            const int d = ref ? (int)(e % ns) : (int)(e / N);
                            const double want = std::fma((double)v[e], s[d], b[d]);
                            const float wantf = (float)((u[e] - b[d]) * rs[d]);
                            ok = std::memcmp(&out[pad + q], &want, 8) == 0 && std::memcmp(&back[q], &wantf, 4) == 0;
                        }
                        if (!ok) {
                            std::printf("FAIL: conversion ns=%d N=%lld ref=%d from %zu, pad %zu\n", ns, N, (int)ref, e0,
                                        pad);
                            ++failures;
                        }
                        ++convs;
                    }
            }
    std::printf("copy crew: %d copies, %d element jobs, %d substrate runs, %d conversions, %d failures\n", copies,
                jobs, walks, convs, failures);
    return failures != 0;
}
