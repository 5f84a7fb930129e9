// biofvm_aux_kernels.h -- gfx950 device code around the sweeps: Thomas
// precompute, agent secretion/uptake, Dirichlet faces, reductions and the
// layout conversions at the host boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "biofvm_kernels.h"

namespace bfa {

// ---- Thomas precompute (FP64), one thread per substrate --------------------
// Constants as BioFVM/BioFVM_solvers.cpp:194-217, pivots and c' as :226-329:
//   c1 = D*dt/dx/dx, c2 = lambda*dt/3, c3 = 1+c1+c1+c2, c3a = 1+c1+c2,
//   den[0] = c3a (1+c2 when n == 1), cp[0] = -c1/den[0],
//   den[i] = (i == n-1 ? c3a : c3) + c1*cp[i-1], cp[i] = -c1/den[i].
// Writes den/cp in FP64 and 1/den, cp in the sweep type T, and kT[d] = c1,
// kT[ns + d] = c2.  For the FP32 deviation form (u = b + s*v) it also writes
// beta = (b/s)*c2: one sweep maps v to T^-1 (v - beta) because every axis
// matrix has unit row sums plus c2.
template <typename T>
__global__ void k_thomas(int n, int ns, const double* __restrict__ params, double dt, double dx,
                         double* __restrict__ den64, double* __restrict__ cp64,
                         T* __restrict__ invT, T* __restrict__ cpT, T* __restrict__ kT,
                         T* __restrict__ betaT) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ns) return;
    const double D = params[d], lam = params[ns + d], b = params[2 * ns + d],
                 s = params[3 * ns + d];
    double c1 = D;
    c1 *= dt;
    c1 /= dx;
    c1 /= dx;
    double c2 = lam;
    c2 *= dt;
    c2 /= 3.0;
    double c3 = 1.0;
    c3 += c1;
    c3 += c1;
    c3 += c2;
    double c3a = 1.0;
    c3a += c1;
    c3a += c2;
    double* den = den64 + (size_t)d * n;
    double* cp = cp64 + (size_t)d * n;
    den[0] = (n == 1) ? 1.0 + c2 : c3a;
    cp[0] = (-c1) / den[0];
    for (int i = 1; i < n; ++i) {
        const double base = (i == n - 1) ? c3a : c3;
        den[i] = fma(c1, cp[i - 1], base);
        cp[i] = (-c1) / den[i];
    }
    for (int i = 0; i < n; ++i) {
        invT[(size_t)d * n + i] = (T)(1.0 / den[i]);
        cpT[(size_t)d * n + i] = (T)cp[i];
    }
    kT[d] = (T)c1;
    kT[ns + d] = (T)c2;
    betaT[d] = (T)((b / s) * c2);
}

// ---- agents: ordered per-voxel composition ----------------------------------
// Agents are sorted by device voxel (stable in agent index) on the host, so
// each segment [off[q], off[q+1]) holds the agents of one voxel in ascending
// agent order.  Each segment's agents are applied in order (agent_segment):
// the reference's serial loop order (BioFVM/BioFVM_microenvironment.cpp:
// 1370-1384) with the per-agent update of BioFVM/BioFVM_basic_agent.cpp:
// 536-552.  FP64: rho = (rho + c1)/c2 + e2 in the reference's operation
// order.  FP32 deviation form: the stored v is lifted to u = b + s*v in FP64,
// updated with the same FP64 operations and stored back as (u - b)/s, so the
// agent maps carry no FP32-rounded coefficients (whose fixed error would be
// amplified near the map's fixed point).
// Constants are stored [ns][nagents_sorted] so lanes of one substrate read
// consecutive entries.
// Internalised-substrate tracking of agent a (sorted position) before its
// update of density rho (track_uptake); trk and e1 are [n][ns] in the
// caller's agent order; orig maps the sorted position to that order.
__device__ __forceinline__ void track_before(double* __restrict__ trk, const int* __restrict__ orig,
                                             int ns, int d, int a, double rho, double c1, double c2,
                                             double W) {
    track_uptake(trk + (size_t)orig[a] * ns + d, rho, c1, c2, W);
}

// ... and after it: I -= e1 (the net export, :553-556).
__device__ __forceinline__ void track_after(double* __restrict__ trk, const int* __restrict__ orig,
                                            const double* __restrict__ e1, int ns, int d, int a) {
    const size_t o = (size_t)orig[a] * ns + d;
    trk[o] = trk[o] - e1[o];
}

// The agents of segment q (one voxel) applied to its density rho of
// substrate d, in ascending agent order: rho = (rho + c1)/c2 + e2 per agent,
// with the tracking updates around each when trk is set.
__device__ __forceinline__ double agent_segment(double rho, int d, int q, const int* __restrict__ off,
                                                const double* __restrict__ c1, const double* __restrict__ c2,
                                                const double* __restrict__ e2, double* __restrict__ trk,
                                                const int* __restrict__ orig, const double* __restrict__ e1,
                                                int ns, int nag, double W) {
    const int a0 = off[q], a1 = off[q + 1];
    const size_t o = (size_t)d * nag;
    for (int a = a0; a < a1; ++a) {
        if (trk) track_before(trk, orig, ns, d, a, rho, c1[o + a], c2[o + a], W);
        rho += c1[o + a];
        rho /= c2[o + a];
        rho += e2[o + a];
        if (trk) track_after(trk, orig, e1, ns, d, a);
    }
    return rho;
}

// One agent step over substrates [d0, d0 + nd), one lane per (segment,
// substrate).  Two ways to store:
//  - LINES == false: each lane reads its voxel, applies its segment and
//    stores the voxel (8 or 4 bytes);
//  - LINES == true: the field is read and written in whole 128-byte lines.
//    The lane of the first segment (in voxel order) of substrate d that falls
//    in a line owns the line: it loads the line, applies every segment of the
//    line in place and stores the whole line back; the lanes of the line's
//    other segments exit.  A line shared with the neighbouring substrate
//    (N*sizeof(T) not a multiple of 128 bytes) is stored voxel by voxel, so
//    each lane writes only its own substrate; the field allocation's zeroed
//    slack covers the last line.
// Single-voxel stores leave lines partly written in the caches, and their
// later write-back (a read-modify-write in memory) can slow the next sweeps
// by more than the agent step takes (at 800^3 x 2 with 10^6 agents, FP64 x
// 3.51 against 3.21 ms); whole lines cost more instructions in the agent
// step.  Which is faster depends on the workload (bfa_set_agent_stores).
// FP64 (params unused) or FP32 deviations lifted to FP64 with the basis
// b, s, 1/s at params[2*ns + d], [3*ns + d], [4*ns + d].
// Segments of more agents than this are applied by k_agents_long, a
// wavefront per segment; k_agents leaves them alone.
constexpr int kLongSegment = 64;

template <typename T, bool LINES>
__global__ void k_agents(T* __restrict__ u, long long N, int ns, int d0, int nd, int nseg, int nag,
                         const int* __restrict__ off, const long long* __restrict__ vox,
                         const double* __restrict__ c1, const double* __restrict__ c2,
                         const double* __restrict__ e2, const double* __restrict__ params,
                         double* __restrict__ trk, const int* __restrict__ orig, const double* __restrict__ e1,
                         double W) {
    constexpr bool DEV = sizeof(T) == 4;
    constexpr int E = 128 / (int)sizeof(T);  // elements per line
    using VA = typename Vec16A<T>::type;
    constexpr int VW = 16 / (int)sizeof(T);
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)nseg * nd) return;
    const int d = d0 + (int)(gid / nseg);
    const int q = (int)(gid - (long long)(d - d0) * nseg);
    const long long lo = (long long)d * N;  // substrate d's elements: [lo, lo + N)
    double b = 0.0, s = 1.0, rs = 1.0;
    if (DEV) {
        b = params[2 * ns + d];
        s = params[3 * ns + d];
        rs = params[4 * ns + d];
    }
    if (!LINES) {
        if (off[q + 1] - off[q] > kLongSegment) return;
        T* p = u + lo + vox[q];
        const double rho = agent_segment(DEV ? fma((double)*p, s, b) : (double)*p, d, q, off, c1, c2, e2, trk,
                                         orig, e1, ns, nag, W);
        *p = DEV ? (T)((rho - b) * rs) : (T)rho;
        return;
    }
    const long long g = (lo + vox[q]) / E;  // the line of this segment's voxel
    if (q > 0 && (lo + vox[q - 1]) / E == g) return;  // an earlier segment's lane owns it
    T* lp = u + g * E;
    T line[E];
#pragma unroll
    for (int i = 0; i < E / VW; ++i) {
        const VA t = reinterpret_cast<const VA*>(lp)[i];
#pragma unroll
        for (int w = 0; w < VW; ++w) line[i * VW + w] = t[w];
    }
    int r = q;
    for (; r < nseg; ++r) {
        const long long v = lo + vox[r];
        if (v / E != g) break;
        if (off[r + 1] - off[r] > kLongSegment) continue;
        const int m = (int)(v - g * E);
        T val = T(0);
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i == m) val = line[i];
        const double rho = agent_segment(DEV ? fma((double)val, s, b) : (double)val, d, r, off, c1, c2, e2,
                                         trk, orig, e1, ns, nag, W);
        val = DEV ? (T)((rho - b) * rs) : (T)rho;
#pragma unroll
        for (int i = 0; i < E; ++i)
            if (i == m) line[i] = val;
    }
    if (g * E >= lo && g * E + E <= lo + N) {
#pragma unroll
        for (int i = 0; i < E / VW; ++i) {
            VA t;
#pragma unroll
            for (int w = 0; w < VW; ++w) t[w] = line[i * VW + w];
            reinterpret_cast<VA*>(lp)[i] = t;
        }
    } else {
        for (int k = q; k < r; ++k) {
            const int m = (int)(lo + vox[k] - g * E);
#pragma unroll
            for (int i = 0; i < E; ++i)
                if (i == m) lp[i] = line[i];
        }
    }
}

// Segments of more than kLongSegment agents (many agents in one voxel), one
// wavefront per (segment, substrate), after k_agents: a segmented reduction
// of the agents' affine maps rho -> rho/c2 + (c1/c2 + e2) over the lanes.
// Lane l takes K = ceil(len/64) consecutive agents and composes their maps;
// an inclusive Hillis-Steele scan of the composed maps over the 64 lanes
// (composition g after f: (A_g A_f, A_g B_f + B_g)) gives each lane the map
// of all agents before its own, so each lane starts its agents from the
// density they see and walks them with the reference's operations
// (rho += c1; rho /= c2; rho += e2, with the tracking updates), and the lane
// holding the segment's last agent stores the voxel.  The lanes' starting
// densities come from composed maps, so the result agrees with the serial
// walk to rounding (a few ulps per lane boundary), not bit for bit.  long_seg
// lists the long segments.  Launch: 64 * nlong * nd threads for substrates
// [d0, d0 + nd).
template <typename T>
__global__ __launch_bounds__(256) void k_agents_long(T* __restrict__ u, long long N, int ns, int d0, int nd,
                                                     int nlong,
                                                     const int* __restrict__ long_seg, int nag,
                                                     const int* __restrict__ off, const long long* __restrict__ vox,
                                                     const double* __restrict__ c1, const double* __restrict__ c2,
                                                     const double* __restrict__ e2,
                                                     const double* __restrict__ params, double* __restrict__ trk,
                                                     const int* __restrict__ orig, const double* __restrict__ e1,
                                                     double W) {
    constexpr bool DEV = sizeof(T) == 4;
    const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wv >= (long long)nlong * nd) return;  // whole wavefronts
    const int d = d0 + (int)(wv / nlong);
    const int q = long_seg[wv - (long long)(d - d0) * nlong];
    const int a0 = off[q], a1 = off[q + 1];
    const int K = (a1 - a0 + 63) / 64;
    const int b0 = min(a1, a0 + lane * K), b1 = min(a1, b0 + K);
    const size_t o = (size_t)d * nag;
    double A = 1.0, B = 0.0;  // this lane's agents composed: rho -> A*rho + B
    for (int a = b0; a < b1; ++a) {
        const double ai = 1.0 / c2[o + a];
        const double bi = c1[o + a] / c2[o + a] + e2[o + a];
        B = fma(ai, B, bi);
        A = ai * A;
    }
#pragma unroll
    for (int sh = 1; sh < 64; sh <<= 1) {
        const double Ap = __shfl_up(A, sh, 64), Bp = __shfl_up(B, sh, 64);
        if (lane >= sh) {
            B = fma(A, Bp, B);
            A = A * Ap;
        }
    }
    // the map of the agents before this lane's: the previous lane's inclusive map
    const double Ae = __shfl_up(A, 1, 64), Be = __shfl_up(B, 1, 64);
    T* p = u + (long long)d * N + vox[q];
    double b = 0.0, s = 1.0, rs = 1.0;
    if (DEV) {
        b = params[2 * ns + d];
        s = params[3 * ns + d];
        rs = params[4 * ns + d];
    }
    const double rho0 = DEV ? fma((double)*p, s, b) : (double)*p;
    double rho = lane == 0 ? rho0 : fma(Ae, rho0, Be);
    for (int a = b0; a < b1; ++a) {
        if (trk) track_before(trk, orig, ns, d, a, rho, c1[o + a], c2[o + a], W);
        rho += c1[o + a];
        rho /= c2[o + a];
        rho += e2[o + a];
        if (trk) track_after(trk, orig, e1, ns, d, a);
    }
    if (b1 == a1 && b0 < b1) *p = DEV ? (T)((rho - b) * rs) : (T)rho;
}

// Appends the index of every segment of more than kLongSegment agents to
// list (in any order) and counts them.
__global__ void k_long_segments(int nseg, const int* __restrict__ off, int* __restrict__ list,
                                int* __restrict__ count) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nseg) return;
    if (off[q + 1] - off[q] > kLongSegment) list[atomicAdd(count, 1)] = q;
}

// ---- standalone Dirichlet pass over the faces -------------------------------
// Thread space: [y faces 2*nz*nx][z faces 2*ny*nx][x faces 2*nz*ny], x fastest
// inside the y and z faces so that their writes are coalesced; yz_only skips
// the x faces (after an x sweep, whose kernel writes the x-end nodes itself).
// Voxels on edges are written more than once, always with the same value.
template <typename T>
__global__ void k_dirichlet_faces(SweepArgs<T> a, int yz_only) {
    const long long fy = (long long)a.nz * a.nx, fz = (long long)a.ny * a.nx, fx = (long long)a.nz * a.ny;
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (g < 2 * fy) {
        j = g >= fy ? a.ny - 1 : 0;
        g -= g >= fy ? fy : 0;
        k = (int)(g / a.nx);
        i = (int)(g - (long long)k * a.nx);
    } else if ((g -= 2 * fy) < 2 * fz) {
        k = g >= fz ? a.nz - 1 : 0;
        g -= g >= fz ? fz : 0;
        j = (int)(g / a.nx);
        i = (int)(g - (long long)j * a.nx);
    } else if (!yz_only && (g -= 2 * fz) < 2 * fx) {
        i = g >= fx ? a.nx - 1 : 0;
        g -= g >= fx ? fx : 0;
        k = (int)(g / a.ny);
        j = (int)(g - (long long)k * a.ny);
    } else {
        return;
    }
    const long long v = ((long long)k * a.ny + j) * a.nx + i;
    for (int d = 0; d < a.ns; ++d)
        if ((a.active >> d) & 1ull) a.u[(long long)d * a.N + v] = dir_value(a, d, i, j, k);
}

// ---- reductions --------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}

// Partial sums of x[off_d + v*stride], v < count, for substrate d = blockIdx.y
// (off_d = d*dstride).  MODE 0: sum of x; MODE 1: max |x - ref[d]|.  Each
// block owns a fixed contiguous slice, so the result does not depend on
// scheduling.  blockDim.x = 256.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_reduce_partial(const T* __restrict__ x, long long count,
                                                         long long stride, long long dstride,
                                                         const double* __restrict__ ref,
                                                         double* __restrict__ partial) {
    __shared__ double sh[4];
    const int d = blockIdx.y;
    const long long per = (count + gridDim.x - 1) / gridDim.x;
    const long long v0 = (long long)blockIdx.x * per;
    const long long v1 = (v0 + per < count) ? v0 + per : count;
    const T* base = x + (long long)d * dstride;
    const double r = MODE == 1 ? ref[d] : 0.0;
    double acc = 0.0;
    for (long long v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
        const double e = (double)base[v * stride];
        if (MODE == 0)
            acc += e;
        else
            acc = fmax(acc, fabs(e - r));
    }
    acc = MODE == 0 ? wave_sum(acc) : wave_max(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = sh[0];
        for (int w = 1; w < 4; ++w) tot = MODE == 0 ? tot + sh[w] : fmax(tot, sh[w]);
        partial[(size_t)d * gridDim.x + blockIdx.x] = tot;
    }
}

// Fixed-order final reduction of nb partials per substrate; one block (64
// threads) per substrate.
template <int MODE>
__global__ void k_reduce_final(const double* __restrict__ partial, int nb,
                               double* __restrict__ out) {
    const int d = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) {
        const double e = partial[(size_t)d * nb + b];
        acc = MODE == 0 ? acc + e : fmax(acc, e);
    }
    acc = MODE == 0 ? wave_sum(acc) : wave_max(acc);
    if (threadIdx.x == 0) out[d] = acc;
}

// ---- host-layout conversion ----------------------------------------------------
// Reference layout in[((i*ny + j)*nz + k)*ns + d] (FP64) to device layout
// out[((d*nz + k)*ny + j)*nx + i] (T), through a 32x32 LDS tile per (j,
// i-block, q-block) with q = k*ns + d, so reads are contiguous in q and
// writes contiguous in i.  v = (u - b[d]) * rs[d] with rs = 1/s (s a power
// of two); FP64 passes b = 0, rs = 1.  blockDim = (32, 8); grid =
// (ceil(nz*ns/32), ceil(nx/32), ny).
template <typename T>
__global__ __launch_bounds__(256) void k_ref_to_soa(const double* __restrict__ in,
                                                     T* __restrict__ out, int nx, int ny,
                                                     int nz, int ns,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ rs) {
    __shared__ double tile[32][33];
    const long long Q = (long long)nz * ns;
    const long long q0 = (long long)blockIdx.x * 32;
    const int i0 = blockIdx.y * 32;
    const int j = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r;
        const long long q = q0 + tx;
        if (i < nx && q < Q) tile[r][tx] = in[((long long)i * ny + j) * Q + q];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const long long q = q0 + r;
        const int i = i0 + tx;
        if (i < nx && q < Q) {
            const int k = (int)(q / ns), d = (int)(q - (long long)k * ns);
            const double v = (tile[tx][r] - b[d]) * rs[d];
            out[(((long long)d * nz + k) * ny + j) * nx + i] = (T)v;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_soa_to_ref(const T* __restrict__ in,
                                                     double* __restrict__ out, int nx, int ny,
                                                     int nz, int ns,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ s) {
    __shared__ double tile[32][33];
    const long long Q = (long long)nz * ns;
    const long long q0 = (long long)blockIdx.x * 32;
    const int i0 = blockIdx.y * 32;
    const int j = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        const long long q = q0 + r;
        const int i = i0 + tx;
        if (i < nx && q < Q) {
            const int k = (int)(q / ns), d = (int)(q - (long long)k * ns);
            tile[r][tx] = fma((double)in[(((long long)d * nz + k) * ny + j) * nx + i], s[d], b[d]);
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r;
        const long long q = q0 + tx;
        if (i < nx && q < Q) out[((long long)i * ny + j) * Q + q] = tile[tx][r];
    }
}

// Device-layout FP64 to T and back, elementwise (substrate d = blockIdx.y).
// The same transposes on FP32 deviations that the host converts itself
// (transfers of FP32 fields at half the PCIe bytes): no arithmetic here.
__global__ __launch_bounds__(256) void k_ref32_to_soa(const float* __restrict__ in, float* __restrict__ out,
                                                      int nx, int ny, int nz, int ns) {
    __shared__ float tile[32][33];
    const long long Q = (long long)nz * ns;
    const long long q0 = (long long)blockIdx.x * 32;
    const int i0 = blockIdx.y * 32;
    const int j = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r;
        const long long q = q0 + tx;
        if (i < nx && q < Q) tile[r][tx] = in[((long long)i * ny + j) * Q + q];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const long long q = q0 + r;
        const int i = i0 + tx;
        if (i < nx && q < Q) {
            const int k = (int)(q / ns), d = (int)(q - (long long)k * ns);
            out[(((long long)d * nz + k) * ny + j) * nx + i] = tile[tx][r];
        }
    }
}

__global__ __launch_bounds__(256) void k_soa_to_ref32(const float* __restrict__ in, float* __restrict__ out,
                                                      int nx, int ny, int nz, int ns) {
    __shared__ float tile[32][33];
    const long long Q = (long long)nz * ns;
    const long long q0 = (long long)blockIdx.x * 32;
    const int i0 = blockIdx.y * 32;
    const int j = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        const long long q = q0 + r;
        const int i = i0 + tx;
        if (i < nx && q < Q) {
            const int k = (int)(q / ns), d = (int)(q - (long long)k * ns);
            tile[r][tx] = in[(((long long)d * nz + k) * ny + j) * nx + i];
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r;
        const long long q = q0 + tx;
        if (i < nx && q < Q) out[((long long)i * ny + j) * Q + q] = tile[tx][r];
    }
}

template <typename T>
__global__ void k_soa_from_f64(const double* __restrict__ in, T* __restrict__ out, long long N,
                               const double* __restrict__ b, const double* __restrict__ rs) {
    const int d = blockIdx.y;
    const double bb = b[d], r = rs[d];
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N;
         v += (long long)gridDim.x * blockDim.x)
        out[(long long)d * N + v] = (T)((in[(long long)d * N + v] - bb) * r);
}

template <typename T>
__global__ void k_soa_to_f64(const T* __restrict__ in, double* __restrict__ out, long long N,
                             const double* __restrict__ b, const double* __restrict__ s) {
    const int d = blockIdx.y;
    const double bb = b[d], ss = s[d];
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N;
         v += (long long)gridDim.x * blockDim.x)
        out[(long long)d * N + v] = fma((double)in[(long long)d * N + v], ss, bb);
}

// Constant fill per substrate: out = (value[d] - b[d]) * rs[d].
template <typename T>
__global__ void k_fill_const(T* __restrict__ out, long long N, const double* __restrict__ value,
                             const double* __restrict__ b, const double* __restrict__ rs) {
    const int d = blockIdx.y;
    const T v = (T)((value[d] - b[d]) * rs[d]);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x)
        out[(long long)d * N + i] = v;
}

// ---- gradient field --------------------------------------------------------
// The derivative of substrate d along one axis at device index v (stride
// between neighbours `st`, position `pos` of n nodes, spacing h): the central
// difference inside, the one-sided ones at the ends, 0 on an axis of one node;
// one subtraction then one division, as compute_all_gradient_vectors
// (BioFVM/BioFVM_microenvironment.cpp:1454-1615).  FP32 fields hold
// u = b + s*v, so du = s*dv.
template <typename T>
__device__ __forceinline__ double grad_axis(const T* __restrict__ ud, long long v, long long st,
                                            int pos, int n, double h, double s) {
    if (n == 1) return 0.0;
    long long hi = v + st, lo = v - st;
    double hh = 2.0 * h;
    if (pos == 0) {
        lo = v;
        hh = h;
    } else if (pos == n - 1) {
        hi = v;
        hh = h;
    }
    if (sizeof(T) == 8) return ((double)ud[hi] - (double)ud[lo]) / hh;
    return (s * ((double)ud[hi] - (double)ud[lo])) / hh;
}

// Whole-field gradient: g[(d*3 + c)*N + v] for device index v.
// Grid: (blocks, ns), grid-stride over v.
template <typename T>
__global__ __launch_bounds__(256) void k_gradient(const T* __restrict__ u, long long N, int nx,
                                                  int ny, int nz, int ns,
                                                  const double* __restrict__ params, double dx,
                                                  double dy, double dz, double* __restrict__ g) {
    const int d = blockIdx.y;
    const T* ud = u + (long long)d * N;
    const double s = sizeof(T) == 8 ? 1.0 : params[3 * ns + d];
    const long long nxy = (long long)nx * ny;
    double* gd = g + (long long)d * 3 * N;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N;
         v += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(v / nxy);
        const long long r = v - (long long)k * nxy;
        const int j = (int)(r / nx);
        const int i = (int)(r - (long long)j * nx);
        gd[v] = grad_axis<T>(ud, v, 1, i, nx, dx, s);
        gd[N + v] = grad_axis<T>(ud, v, nx, j, ny, dy, s);
        gd[2 * N + v] = grad_axis<T>(ud, v, nxy, k, nz, dz, s);
    }
}

// Gradient at listed voxels (reference voxel index n = (i*ny + j)*nz + k):
// out[(a*ns + d)*3 + c].  One lane per (voxel, substrate).
template <typename T>
__global__ __launch_bounds__(256) void k_gradient_at(const T* __restrict__ u, long long N, int nx,
                                                     int ny, int nz, int ns,
                                                     const double* __restrict__ params, double dx,
                                                     double dy, double dz, int n,
                                                     const int* __restrict__ voxel,
                                                     double* __restrict__ out) {
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (long long)n * ns) return;
    const int a = (int)(w / ns);
    const int d = (int)(w - (long long)a * ns);
    const int vr = voxel[a];
    const int k = vr % nz;
    const int j = (vr / nz) % ny;
    const int i = vr / (nz * ny);
    const long long nxy = (long long)nx * ny;
    const long long v = (long long)k * nxy + (long long)j * nx + i;
    const T* ud = u + (long long)d * N;
    const double s = sizeof(T) == 8 ? 1.0 : params[3 * ns + d];
    double* o = out + w * 3;
    o[0] = grad_axis<T>(ud, v, 1, i, nx, dx, s);
    o[1] = grad_axis<T>(ud, v, nx, j, ny, dy, s);
    o[2] = grad_axis<T>(ud, v, nxy, k, nz, dz, s);
}

// ---- bulk sources and sinks ---------------------------------------------------
// rho <- (rho + dt*ST) / ((U+S)*dt + 1) per value, with ST = T*S and
// UpS = U + S precomputed in FP64 on the host, in the operation order of
// simulate_bulk_sources_and_sinks (BioFVM/BioFVM_microenvironment.cpp:1327-1368):
// one fused multiply-add for rho += dt*t2, separate roundings for t3*dt and
// t3 + 1 (floating-point contraction is switched off in the kernel).  FP32 fields lift v to u = b + s*v in FP64 and
// store (u - b)/s.  Grid: (blocks, ns).
template <typename T>
__global__ __launch_bounds__(256) void k_bulk(T* __restrict__ u, long long N, int ns,
                                              const double* __restrict__ params,
                                              const double* __restrict__ st,
                                              const double* __restrict__ ups, double dt) {
#pragma clang fp contract(off)
    const int d = blockIdx.y;
    const bool dev = sizeof(T) == 4;
    const double b = dev ? params[2 * ns + d] : 0.0;
    const double s = dev ? params[3 * ns + d] : 1.0;
    const double rs = dev ? params[4 * ns + d] : 1.0;
    T* ud = u + (long long)d * N;
    const double* std_ = st + (long long)d * N;
    const double* upd = ups + (long long)d * N;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < N;
         v += (long long)gridDim.x * blockDim.x) {
        const double x = dev ? fma((double)ud[v], s, b) : (double)ud[v];
        const double rho = fma(dt, std_[v], x);
        const double den = upd[v] * dt + 1.0;  // two roundings: contraction is off here
        const double y = rho / den;
        ud[v] = dev ? (T)((y - b) * rs) : (T)y;
    }
}

// Plain 16-byte copy, used only to calibrate HBM byte counters.
__global__ void k_copy16(const uint4* __restrict__ in, uint4* __restrict__ out, long long n) {
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x)
        out[v] = in[v];
}

// One-pass streaming copy, the achievable-bandwidth reference: the grid
// covers the buffer once, each thread moving four 16-byte vectors (a block's
// 256 threads apart) with nontemporal loads and stores.  Measured on MI355X
// at 1 GiB: 6.0-6.3 TB/s, against 5.4-5.6 TB/s for k_copy16's grid-stride
// loop over 65536 blocks.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_stream(const u32x4* __restrict__ in,
                                                     u32x4* __restrict__ out, long long n) {
    constexpr int U = 4;
    const long long b = (long long)blockIdx.x * blockDim.x * U + threadIdx.x;
    u32x4 t[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const long long v = b + (long long)q * blockDim.x;
        if (v < n) t[q] = __builtin_nontemporal_load(in + v);
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const long long v = b + (long long)q * blockDim.x;
        if (v < n) __builtin_nontemporal_store(t[q], out + v);
    }
}

}  // namespace bfa
