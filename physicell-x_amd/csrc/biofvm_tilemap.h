// biofvm_tilemap.h -- workgroup-to-tile maps of the sweep launches.
//
// The workgroups of a launch are dispatched round-robin over the 8 XCDs
// (linear workgroup id mod 8), each XCD with its own L2.  These maps turn the
// linear id b of G workgroups into a logical tile index; both are bijections
// on [0, G).  They are plain constexpr integer functions, callable from device
// code and from host-side tests alike.
#pragma once

namespace bfa {

// Every XCD runs one contiguous run of tiles, in order: XCD x takes the
// q + (x < r) tiles starting at x*q + min(x, r) (q = G / 8, r = G % 8), its
// i-th workgroup tile i of that run.  Neighbouring tiles that share cache
// lines at their edges then meet in the same L2.
constexpr int xcd_block(int b, int G) {
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// The same per-XCD share cut into R sub-runs of consecutive tiles (lengths
// differing by at most one, the longer ones first) that the XCD advances
// together: its i-th workgroup takes position i / R of sub-run i % R.  The
// workgroups an XCD holds at one time then spread over R separate stretches
// of the field instead of one.  R = 1 is xcd_block; R >= 1.
constexpr int xcd_runs(int b, int G, int R) {
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, i = b >> 3;
    const int len = q + (x < r ? 1 : 0);
    const int start = x * q + (x < r ? x : r);
    const int m = len / R, rr = len - m * R;
    const int sub = i % R, pos = i / R;
    return start + sub * m + (sub < rr ? sub : rr) + pos;
}

}  // namespace bfa
