// biofvm_kernels.h -- gfx950 device code for the BioFVM LOD diffusion-decay step.
//
// Device layout (one microenvironment): the field is substrate-major SoA,
//   u[((d*nz + k)*ny + j)*nx + i],
// so x is the unit-stride axis.  One LOD step is three sweeps, each solving
// the constant-coefficient tridiagonal system of every grid line along one
// axis (BioFVM/BioFVM_solvers.cpp:352-777), with the reference's four
// Dirichlet passes (:345,682,731,778) fused into the sweeps: the first and
// second into the x sweep's prologue and epilogue, the third into the y
// sweep's epilogue, the fourth into the z sweep's epilogue.
//
// Line solve.  For a line r[0..n-1] the reference computes
//   forward   y[0] = r[0]/den[0],  y[i] = (r[i] + k*y[i-1])/den[i]
//   backward  x[n-1] = y[n-1],     x[i] = y[i] - cp[i]*x[i+1]
// with den/cp the Thomas LU pivots and c' coefficients.  Here a line is cut
// into P chunks of L consecutive nodes, one chunk per thread, held in
// registers.  Both recurrences are first-order linear, so a chunk's result is
// its zero-carry result plus a carry-in times the product of its
// multipliers.  Each thread computes (zero-carry end value, multiplier
// product) for its chunk, the P pairs are exchanged through LDS, every thread
// folds the pairs of the chunks before it (forward) or after it (backward) to
// obtain its exact carry-in, then reruns the chunk recurrence from that
// carry.  Every field value is read from HBM once and written once per sweep.
//
// Loads are coalesced on every axis: the y and z sweeps give each lane one
// x-column (64 adjacent columns per wavefront), and the x sweep stages a tile
// of R whole rows through LDS with an odd row pitch, so lanes that own the
// same chunk of different rows hit distinct banks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "biofvm_tilemap.h"

namespace bfa {

// Face selection for a boundary voxel: the last registered face (order xmin,
// xmax, ymin, ymax, zmin, zmax) containing it, else slot 6 (value 1.0).
// Mirrors the add_dirichlet_node replay of
// BioFVM/BioFVM_microenvironment.cpp:218-226,2649-2786.
__device__ __forceinline__ int dir_slot(int i, int j, int k, int nx, int ny, int nz,
                                        unsigned reg) {
    if (k == nz - 1 && (reg & 32u)) return 5;
    if (k == 0 && (reg & 16u)) return 4;
    if (j == ny - 1 && (reg & 8u)) return 3;
    if (j == 0 && (reg & 4u)) return 2;
    if (i == nx - 1 && (reg & 2u)) return 1;
    if (i == 0 && (reg & 1u)) return 0;
    return 6;
}

template <typename T>
struct SweepArgs {
    T* u;             // field base of this microenvironment (SoA)
    const T* inv;     // [ns][n] 1/den for the swept axis
    const T* cp;      // [ns][n] c' for the swept axis
    const T* k;       // [2][ns]: D*dt/dx^2, then decay*dt/3
    const T* beta;    // [ns] deviation shift (FP32 form), unused for FP64
    const T* dval;    // [7][ns] Dirichlet value per face slot
    // Per-voxel Dirichlet values of the six face planes, or nullptr to use
    // dval; read only by the standalone face pass (k_dirichlet_faces): with
    // planes the sweeps apply no Dirichlet values themselves and the face
    // pass runs between them.  Substrate d's planes start at plane + d*shell,
    // in the order
    //   xmin [k][j], xmax [k][j]   (ny*nz each)
    //   ymin [k][i], ymax [k][i]   (nx*nz each)
    //   zmin [j][i], zmax [j][i]   (nx*ny each)
    // A voxel on several faces holds the same value in each of their planes.
    const T* plane;
    long long shell;  // 2*(ny*nz + nx*nz + nx*ny)
    unsigned long long active;  // bit d set: substrate d has Dirichlet on
    long long N;      // nx*ny*nz
    int nx, ny, nz, ns;
    unsigned reg;     // bit f set: face f registered
    // Launch window: substrates d0 + blockIdx.y; lines [first, first + count)
    // of each substrate (x sweep: rows r = k*ny + j; y sweep: columns
    // c = k*nx + i; z sweep: columns c = j*nx + i).  A k-slab [k0, k1) is the
    // window first = k0*ny (x) or k0*nx (y), count = (k1-k0)*ny or (k1-k0)*nx.
    int d0;
    long long first, count;
    // Column sweeps: sub-runs of tiles each XCD advances together (xcd_runs;
    // 1: one contiguous run per XCD).  Unused by the x sweeps.
    int runs;
};

// Offsets of the six face planes inside one substrate's shell.
struct PlaneOffsets {
    long long xlo, xhi, ylo, yhi, zlo, zhi;
};

template <typename T>
__device__ __forceinline__ PlaneOffsets plane_offsets(const SweepArgs<T>& a) {
    const long long nyz = (long long)a.ny * a.nz, nxz = (long long)a.nx * a.nz,
                    nxy = (long long)a.nx * a.ny;
    return PlaneOffsets{0, nyz, 2 * nyz, 2 * nyz + nxz, 2 * (nyz + nxz), 2 * (nyz + nxz) + nxy};
}

// Dirichlet value of boundary voxel (i, j, k) of substrate d: the per-voxel
// plane value when planes are set (any plane holding the voxel: they agree),
// else the face table entry chosen by dir_slot.
template <typename T>
__device__ __forceinline__ T dir_value(const SweepArgs<T>& a, int d, int i, int j, int k) {
    if (a.plane) {
        const T* pl = a.plane + (long long)d * a.shell;
        const PlaneOffsets o = plane_offsets(a);
        if (i == 0) return pl[o.xlo + (long long)k * a.ny + j];
        if (i == a.nx - 1) return pl[o.xhi + (long long)k * a.ny + j];
        if (j == 0) return pl[o.ylo + (long long)k * a.nx + i];
        if (j == a.ny - 1) return pl[o.yhi + (long long)k * a.nx + i];
        if (k == 0) return pl[o.zlo + (long long)j * a.nx + i];
        return pl[o.zhi + (long long)j * a.nx + i];
    }
    return a.dval[dir_slot(i, j, k, a.nx, a.ny, a.nz, a.reg) * a.ns + d];
}

// Internalised-substrate tracking of one agent on its total I, before the
// agent's update of density rho (track_internalized_substrates_in_each_agent,
// BioFVM/BioFVM_basic_agent.cpp:520-532): I -= W*((1 - c2)*rho + c1)/c2,
// each operation rounded on its own as the reference's vector statements.
__device__ __forceinline__ void track_uptake(double* I, double rho, double c1, double c2, double W) {
#pragma clang fp contract(off)
    double t = 1.0;
    t -= c2;
    t *= rho;
    t += c1;
    t /= c2;
    t *= W;
    *I = *I - t;
}

// ---- chunk recurrences (positions s..s+L-1 of a padded line) ----------------
// Coefficient arrays are padded to P*L entries with 1/den = c' = 0 past the
// line end, and padded values are 0, so the recurrences need no masking: the
// forward pass carries zeros through the padding and the backward pass starts
// from zero at the true line end.

template <typename T, int L>
__device__ __forceinline__ void fwd_local(const T (&x)[L], int s, const T* __restrict__ inv, T k,
                                          T& zend, T& G) {
    T z = T(0), g = T(1);
#pragma unroll
    for (int m = 0; m < L; ++m) {
        const T iv = inv[s + m];
        z = fma(k, z, x[m]) * iv;
        g = g * (k * iv);
    }
    zend = z;
    G = g;
}

template <typename T, int L>
__device__ __forceinline__ void fwd_apply(T (&x)[L], int s, const T* __restrict__ inv, T k, T y) {
#pragma unroll
    for (int m = 0; m < L; ++m) {
        y = fma(k, y, x[m]) * inv[s + m];
        x[m] = y;
    }
}

template <typename T, int L>
__device__ __forceinline__ void bwd_local(const T (&x)[L], int s, const T* __restrict__ cp,
                                          T& wstart, T& H) {
    T w = T(0), h = T(1);
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
        const T c = cp[s + m];
        w = fma(-c, w, x[m]);
        h = h * (-c);
    }
    wstart = w;
    H = h;
}

template <typename T, int L>
__device__ __forceinline__ void bwd_apply(T (&x)[L], int s, const T* __restrict__ cp, T xn) {
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
        xn = fma(-cp[s + m], xn, x[m]);
        x[m] = xn;
    }
}

// Full chunked solve of one line chunk held in x[].  s_z holds one carry per
// (chunk, line) at s_z[p*lines + line]; s_g one multiplier product per chunk.
// All threads of the block must call this (it contains barriers); threads
// with act == false only take part in the barriers.
template <typename T, int L>
__device__ __forceinline__ void chunk_solve(T (&x)[L], bool act, int p, int P, int line,
                                            int lines, const T* __restrict__ s_inv,
                                            const T* __restrict__ s_cp, T k, T* s_z, T* s_g) {
    const int s = p * L;
    T a = T(0), g = T(1);
    if (act) {
        fwd_local<T, L>(x, s, s_inv, k, a, g);
        s_z[p * lines + line] = a;
        if (line == 0) s_g[p] = g;
    }
    __syncthreads();
    if (act) {
        T y = T(0);
#pragma unroll 8
        for (int q = 0; q < p; ++q) y = fma(s_g[q], y, s_z[q * lines + line]);
        fwd_apply<T, L>(x, s, s_inv, k, y);
    }
    __syncthreads();
    if (act) {
        bwd_local<T, L>(x, s, s_cp, a, g);
        s_z[p * lines + line] = a;
        if (line == 0) s_g[p] = g;
    }
    __syncthreads();
    if (act) {
        T xn = T(0);
#pragma unroll 8
        for (int q = P - 1; q > p; --q) xn = fma(s_g[q], xn, s_z[q * lines + line]);
        bwd_apply<T, L>(x, s, s_cp, xn);
    }
}

// Stage padded coefficients of substrate d into LDS: s_inv/s_cp[0..PL).
template <typename T>
__device__ __forceinline__ void stage_coefs(T* s_inv, T* s_cp, const T* __restrict__ inv,
                                            const T* __restrict__ cp, int n, int PL, int t,
                                            int nthr) {
    for (int i = t; i < PL; i += nthr) {
        s_inv[i] = i < n ? inv[i] : T(0);
        s_cp[i] = i < n ? cp[i] : T(0);
    }
}

// ---- FP32 scaled-deviation form: the line solve in delta form --------------
// One sweep maps v to x = T^-1 (v - beta).  Written as x = v + delta with
//   T delta = r,  r = (v - beta) - T v
//     r_i = -beta - lam*v_i + k*((v_{i-1} - v_i) + (v_{i+1} - v_i))
// (no neighbour term past either end of the line: Neumann rows), the FP32
// arithmetic acts on the small residual and increment instead of the values
// themselves.  The plain form rounds T^-1 (v - beta) in FP32 with
// FP32-rounded coefficients, which perturbs the operator's decay factor by a
// few units in the last place; near a steady state that shifts the fixed
// point by about eps32 / (decay * dt) relative (1e-5 after a few hundred
// steps at decay 0.35 /min).  In delta form the fixed point is where r = 0,
// set by k, lam and beta alone, and the coefficient rounding only changes how
// fast delta converges.
// x[] holds the chunk's values at nodes s..s+L-1 of a line of n nodes, vl and
// vr those at s-1 and s+L (ignored past the line ends); x[] is overwritten
// with the residual, zero past the line end.  The values themselves are kept
// by the caller (in LDS) for x = v + delta after the solve.
// Each link difference v[i+1] - v[i] is formed once and serves both nodes
// (lap_i = d_i - d_{i-1}, with d = 0 across a line end).
template <typename T, int L>
__device__ __forceinline__ void delta_residual(T (&x)[L], T vl, T vr, int s, int n, T k, T lam, T beta) {
    T dprev = s > 0 ? x[0] - vl : T(0);
#pragma unroll
    for (int m = 0; m < L; ++m) {
        const int pos = s + m;
        const T v = x[m];
        const T right = m == L - 1 ? vr : x[m + 1];
        const T dnext = pos < n - 1 ? right - v : T(0);
        x[m] = pos < n ? fma(k, dnext - dprev, fma(-lam, v, -beta)) : T(0);
        dprev = dnext;
    }
}

// A column chunk's values in LDS, node-major with the TX columns of a block
// side by side (conflict-free): node m of chunk p, column lane at
// s_v[(p*L + m)*TX + lane].
template <typename T, int L>
__device__ __forceinline__ void col_stash(T* s_v, const T (&x)[L], int p, int lane, int TX) {
#pragma unroll
    for (int m = 0; m < L; ++m) s_v[(p * L + m) * TX + lane] = x[m];
}

template <typename T, int L>
__device__ __forceinline__ void col_unstash_add(const T* s_v, T (&x)[L], int p, int lane, int TX) {
#pragma unroll
    for (int m = 0; m < L; ++m) x[m] = s_v[(p * L + m) * TX + lane] + x[m];
}

// ---- x sweep: R rows per block staged through LDS ---------------------------
// blockDim.x = roundup(R*P, 64) with P = ceil(nx/L); grid = (ceil(ny*nz/R), ns).
// Dynamic LDS: (R*S + 2*P*L + P*R + P) * sizeof(T), S = nx | 1.
template <typename T, int L, bool DEV>
__global__ __launch_bounds__(1024) void k_sweep_x(SweepArgs<T> a, int R, int logR) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int P = (nx + L - 1) / L;
    const int S = nx | 1;
    T* tile = reinterpret_cast<T*>(smem);
    T* s_inv = tile + (size_t)R * S;
    T* s_cp = s_inv + P * L;
    T* s_z = s_cp + P * L;
    T* s_g = s_z + (size_t)P * R;

    const int t = threadIdx.x;
    const int nthr = blockDim.x;
    const int d = blockIdx.y;
    const long long rows = (long long)ny * nz;
    const long long r0 = (long long)blockIdx.x * R;
    const int rvalid = (int)((rows - r0) < R ? (rows - r0) : R);
    T* g = a.u + (long long)d * a.N + r0 * nx;

    stage_coefs<T>(s_inv, s_cp, a.inv + (long long)d * nx, a.cp + (long long)d * nx, nx, P * L, t,
                   nthr);
    const int wave = t >> 6, lane = t & 63, nw = nthr >> 6;
    for (int row = wave; row < rvalid; row += nw) {
        const T* src = g + (long long)row * nx;
        T* dst = tile + row * S;
        for (int col = lane; col < nx; col += 64) dst[col] = src[col];
    }
    __syncthreads();

    const int rho = t & (R - 1);
    const int p = t >> logR;
    const bool act = (p < P) && (rho < rvalid);
    const int s = p * L;
    int jj = 0, kk = 0;
    bool rowb = false;
    if (act) {
        const long long r = r0 + rho;
        kk = (int)(r / ny);
        jj = (int)(r - (long long)kk * ny);
        rowb = (jj == 0 || jj == ny - 1 || kk == 0 || kk == nz - 1);
    }
    const bool dact = (a.active >> d) & 1ull;
    T x[L];
#pragma unroll
    for (int m = 0; m < L; ++m) x[m] = (act && s + m < nx) ? tile[rho * S + s + m] : T(0);
    if (act && dact) {
#pragma unroll
        for (int m = 0; m < L; ++m) {
            const int pos = s + m;
            if (pos < nx && (rowb || pos == 0 || pos == nx - 1))
                x[m] = a.dval[dir_slot(pos, jj, kk, nx, ny, nz, a.reg) * a.ns + d];
        }
    }
    if (DEV) {
        // the neighbours of the chunk's end nodes, after the Dirichlet pass;
        // the tile keeps v for x = v + delta
        if (act) {
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (s + m < nx) tile[rho * S + s + m] = x[m];
        }
        __syncthreads();
        const T vl = (act && s > 0) ? tile[rho * S + s - 1] : T(0);
        const T vr = (act && s + L < nx) ? tile[rho * S + s + L] : T(0);
        delta_residual<T, L>(x, vl, vr, s, act ? nx : 0, a.k[d], a.k[a.ns + d], a.beta[d]);
        chunk_solve<T, L>(x, act, p, P, rho, R, s_inv, s_cp, a.k[d], s_z, s_g);
        if (act) {
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (s + m < nx) x[m] = tile[rho * S + s + m] + x[m];
        }
    } else {
        chunk_solve<T, L>(x, act, p, P, rho, R, s_inv, s_cp, a.k[d], s_z, s_g);
    }
    if (act && dact) {
#pragma unroll
        for (int m = 0; m < L; ++m) {
            const int pos = s + m;
            if (pos < nx && (rowb || pos == 0 || pos == nx - 1))
                x[m] = a.dval[dir_slot(pos, jj, kk, nx, ny, nz, a.reg) * a.ns + d];
        }
    }
    if (act) {
#pragma unroll
        for (int m = 0; m < L; ++m)
            if (s + m < nx) tile[rho * S + s + m] = x[m];
    }
    __syncthreads();
    for (int row = wave; row < rvalid; row += nw) {
        T* dst = g + (long long)row * nx;
        const T* src = tile + row * S;
        for (int col = lane; col < nx; col += 64) dst[col] = src[col];
    }
}

// ---- x sweep, wave-local variant -------------------------------------------
// A row of nx nodes is cut into P = ceil(nx/L) chunks owned by P consecutive
// lanes of one wavefront, so a wave holds RW = 64/P whole rows and never
// synchronises with other waves.  Each lane reads its L contiguous nodes with
// 16-byte vector loads (at any element alignment), the chunk carries are
// combined with a Hillis-Steele scan of affine maps y -> a*y + b over the
// lanes of the row (shuffles, log2(P) steps each way), and rows are
// prefetched one group ahead so that loads of group g+1 overlap the solve of
// group g.  A solved group is stored through a per-wave LDS buffer, so each
// store instruction writes 64 consecutive vectors (x_row_solve_staged).  The
// lane's 1/den and c' values stay in registers for the whole kernel.  Grid:
// (blocks, ns) with 256-thread blocks; each wave walks row groups
// g = wave_id, wave_id + total_waves, ...
// 16-byte vectors with only the element's alignment: a row chunk starts at
// an arbitrary element, and gfx950 takes unaligned global vector accesses.
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = f32x4u;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    using type = f64x2u;
    static constexpr int n = 2;
};
// the same vectors at their natural 16-byte alignment (LDS staging)
typedef float f32x4a __attribute__((ext_vector_type(4)));
typedef double f64x2a __attribute__((ext_vector_type(2)));
template <typename T>
struct Vec16A;
template <>
struct Vec16A<float> {
    using type = f32x4a;
};
template <>
struct Vec16A<double> {
    using type = f64x2a;
};

// Load the L nodes of a row chunk with 16-byte vector loads.  Nodes past the
// row end are read too (the next row's, or the zeroed slack that follows
// every field allocation) and are ignored by the solve.
template <typename T, int L>
__device__ __forceinline__ void load_chunk(T (&x)[L], const T* __restrict__ src) {
    using V = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::n;
    static_assert(L % W == 0, "chunk length must fill 16-byte vectors");
    const V* v = reinterpret_cast<const V*>(src);
#pragma unroll
    for (int q = 0; q < L / W; ++q) {
        // A lane's chunk of 64 bytes or more spans (most of) a 128-byte line
        // that the lane's successive 16-byte loads fetch one after another;
        // streaming loads let the line leave L2 before the last of them and
        // fetch it again (FP64 chunks of 16 at nx = 800 read 1.42x the field,
        // 2.96 TB/s; plain loads 4.99 TB/s).  Shorter chunks stream.
        const V t = L * sizeof(T) >= 64 ? v[q] : __builtin_nontemporal_load(v + q);
#pragma unroll
        for (int w = 0; w < W; ++w) x[q * W + w] = t[w];
    }
}

// Overwrite the Dirichlet nodes of the chunk [s, s+L) of a row of n nodes:
// the ends take v_lo / v_hi, and when the whole row is on a face (rowb) the
// interior takes v_in.  Static register indices only.  Nodes past the row end
// take v_in too on a face row: no caller reads or stores them (the solves run
// them on zero coefficients and the stores stop at the row end), and one
// position test per node remains instead of three.
template <typename T, int L>
__device__ __forceinline__ void dirichlet_row(T (&x)[L], int s, int n, bool rowb, T v_lo,
                                              T v_hi, T v_in) {
    if (rowb) {
#pragma unroll
        for (int m = 0; m < L; ++m) x[m] = v_in;
    }
    if (s == 0) x[0] = v_lo;
#pragma unroll
    for (int m = 0; m < L; ++m)
        if (s + m == n - 1) x[m] = v_hi;
}

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int d) { return __shfl_up(v, d, 64); }
template <typename T>
__device__ __forceinline__ T shfl_down_t(T v, int d) { return __shfl_down(v, d, 64); }

// ---- wave-local x sweep: per-lane state and row routines ------------------
// A row's P = ceil(nx/L) chunks sit in P consecutive lanes of a wavefront
// (RW = 64/P rows per wavefront); lane p owns nodes [p*L, p*L+L).  The chunk
// carries are combined by a Hillis-Steele scan of the affine maps
// y -> A*y + B over the lanes of the row.  The A parts (products of the
// multipliers k/den and -c') depend on the coefficients only, so every
// lane's scan multipliers are computed once (XCoef) and each row's scan
// shuffles only the B parts.

// Dirichlet face slot of an x-end node of a row that is on no y or z face.
__device__ __forceinline__ int dir_slot_xonly(int i, int nx, unsigned reg) {
    if (i == nx - 1 && (reg & 2u)) return 1;
    if (i == 0 && (reg & 1u)) return 0;
    return 6;
}

template <typename T, int L>
struct XCoef {
    T inv[L], cp[L];  // 1/den and c' of the lane's nodes, zero past the row end
    T fm[6], bm[6];   // multiplier of forward / backward scan step j, 0 where the step is off
    T k, bt, lam;     // D*dt/dx^2, the deviation shift and decay*dt/3
    T v_lo, v_hi;     // Dirichlet values at the x ends of rows on no y/z face
    bool dact;        // Dirichlet on for this substrate
};

template <typename T, int L, bool DEV>
__device__ __forceinline__ void x_coef_setup(const SweepArgs<T>& a, int d, int p, int P, int s,
                                             int valid, XCoef<T, L>& c) {
    const int nx = a.nx;
#pragma unroll
    for (int m = 0; m < L; ++m) {
        c.inv[m] = m < valid ? a.inv[d * nx + s + m] : T(0);
        c.cp[m] = m < valid ? a.cp[d * nx + s + m] : T(0);
    }
    c.k = a.k[d];
    c.bt = DEV ? a.beta[d] : T(0);
    c.lam = DEV ? a.k[a.ns + d] : T(0);
    c.dact = (a.active >> d) & 1ull;
    c.v_lo = a.dval[dir_slot_xonly(0, nx, a.reg) * a.ns + d];
    c.v_hi = a.dval[dir_slot_xonly(nx - 1, nx, a.reg) * a.ns + d];
    T za = T(1);
#pragma unroll
    for (int m = 0; m < L; ++m) za = za * (c.k * c.inv[m]);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int off = 1 << j;
        const T pa = shfl_up_t(za, off);
        const bool on = off < P && p >= off;
        c.fm[j] = on ? za : T(0);
        if (on) za = za * pa;
    }
    T wa = T(1);
#pragma unroll
    for (int m = L - 1; m >= 0; --m) wa = wa * (-c.cp[m]);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int off = 1 << j;
        const T na = shfl_down_t(wa, off);
        const bool on = off < P && p + off < P;
        c.bm[j] = on ? wa : T(0);
        if (on) wa = wa * na;
    }
}

// One row group.  x_row_fetch issues the row's loads (the field chunk, and
// the per-row Dirichlet values of a row on a y or z face); x_row_solve_staged
// consumes them, so a caller can fetch the next row group before solving the
// current one.  Dirichlet passes D0 (before) and D1 (after the solve) set
// the row's end nodes, and all its nodes when the row lies on a y or z face.
template <typename T, int L>
struct XRow {
    T v[L];
    T f_lo, f_hi, f_in;  // Dirichlet values of a row on a y or z face
    bool rowb;           // the row is on a y or z face and Dirichlet is on
};

// The per-row Dirichlet part of x_row_fetch: rowb and the face values of
// row r (= k*ny + j) when it lies on a y or z face.  FR = false leaves rows on
// y and z faces to a face pass after the sweep (k_dirichlet_faces): every row
// then takes only its two x-end values, from registers, which keeps the
// per-row face lookups out of the kernel (rows are solved independently, so
// a face row's result only has to be overwritten before the next sweep).
template <typename T, int L, bool FR = true>
__device__ __forceinline__ void x_row_dir_fetch(const SweepArgs<T>& a, int d, int r, bool have,
                                                const XCoef<T, L>& c, XRow<T, L>& q) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    q.f_lo = q.f_hi = q.f_in = T(0);
    q.rowb = false;
    if (FR && have && c.dact) {
        const int kk = r / ny;
        const int jj = r - kk * ny;
        q.rowb = (jj == 0 || jj == ny - 1 || kk == 0 || kk == nz - 1);
        if (q.rowb) {
            q.f_lo = a.dval[dir_slot(0, jj, kk, nx, ny, nz, a.reg) * a.ns + d];
            q.f_hi = a.dval[dir_slot(nx - 1, jj, kk, nx, ny, nz, a.reg) * a.ns + d];
            q.f_in = a.dval[dir_slot(-1, jj, kk, nx, ny, nz, a.reg) * a.ns + d];
        }
    }
}

template <typename T, int L, bool FR = true>
__device__ __forceinline__ void x_row_fetch(const SweepArgs<T>& a, int d, int r, bool have, int s,
                                            int valid, const XCoef<T, L>& c, XRow<T, L>& q) {
    const T* ub = a.u + (long long)d * a.N;
    if (have)
        load_chunk<T, L>(q.v, ub + (long long)r * a.nx + s);
    else {
#pragma unroll
        for (int m = 0; m < L; ++m) q.v[m] = T(0);
    }
    x_row_dir_fetch<T, L, FR>(a, d, r, have, c, q);
}

// Dirichlet pass of one row chunk (see x_row_fetch).
template <typename T, int L>
__device__ __forceinline__ void x_row_dirichlet(T (&cur)[L], int nx, int s, bool have,
                                                const XCoef<T, L>& c, const XRow<T, L>& q) {
    if (q.rowb) {
        dirichlet_row<T, L>(cur, s, nx, true, q.f_lo, q.f_hi, q.f_in);
    } else if (have && c.dact) {
        if (s == 0) cur[0] = c.v_lo;
        if (nx % L == 0) {
            if (s + L == nx) cur[L - 1] = c.v_hi;
        } else {
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (s + m == nx - 1) cur[m] = c.v_hi;
        }
    }
}

// Dirichlet D0, the row solve and Dirichlet D1 on the chunk held in q.v.
// FP64 solves in place on the values.  FP32 (DEV) solves for the increment
// from the residual (delta_residual): the neighbours of the chunk's end nodes
// are the last node of lane p-1 and the first of lane p+1 (lanes of other
// rows are masked by the line-end tests), and after the solve the values are
// read back from vs, this lane's chunk in LDS (valid nodes), for v + delta;
// vs holds v at every node that is not a Dirichlet node (D1 rewrites those).
template <typename T, int L, bool DEV>
__device__ __forceinline__ void x_row_compute(int nx, bool have, int p, int P, int s, int valid,
                                              const XCoef<T, L>& c, XRow<T, L>& q, const T* vs) {
    x_row_dirichlet<T, L>(q.v, nx, s, have, c, q);
    if (DEV) {
        const T vl = shfl_up_t(q.v[L - 1], 1);
        const T vr = shfl_down_t(q.v[0], 1);
        delta_residual<T, L>(q.v, vl, vr, s, nx, c.k, c.lam, c.bt);
    }
    T(&cur)[L] = q.v;
    // past the row end 1/den = c' = 0, so whatever sits there leaves the
    // row's values unchanged and becomes 0 in the forward pass
    T zb = T(0);
#pragma unroll
    for (int m = 0; m < L; ++m) zb = fma(c.k, zb, cur[m]) * c.inv[m];
    // scan steps that are off have multiplier 0 (and a shuffle source inside
    // the wave), so all six run unconditionally
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int off = 1 << j;
        const T pb = shfl_up_t(zb, off);
        zb = fma(c.fm[j], pb, zb);
    }
    T y = shfl_up_t(zb, 1);
    if (p == 0) y = T(0);
#pragma unroll
    for (int m = 0; m < L; ++m) {
        y = fma(c.k, y, cur[m]) * c.inv[m];
        cur[m] = y;
    }
    T wb = T(0);
#pragma unroll
    for (int m = L - 1; m >= 0; --m) wb = fma(-c.cp[m], wb, cur[m]);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int off = 1 << j;
        const T nb = shfl_down_t(wb, off);
        wb = fma(c.bm[j], nb, wb);
    }
    T xn = shfl_down_t(wb, 1);
    if (p == P - 1) xn = T(0);
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
        xn = fma(-c.cp[m], xn, cur[m]);
        cur[m] = xn;
    }
    if (DEV && have) {
        using VA = typename Vec16A<T>::type;
        constexpr int W = Vec16<T>::n;
        if (L % W == 0 && nx % W == 0) {  // vs is vector aligned and valid a multiple of W
#pragma unroll
            for (int v = 0; v < L / W; ++v)
                if (v * W < valid) {
                    const VA t = *reinterpret_cast<const VA*>(vs + v * W);
#pragma unroll
                    for (int w = 0; w < W; ++w) q.v[v * W + w] = t[w] + q.v[v * W + w];
                }
        } else {
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (m < valid) q.v[m] = vs[m] + q.v[m];
        }
    }
    if (have) x_row_dirichlet<T, L>(q.v, nx, s, have, c, q);
}

// Store of a solved row group through LDS.  A row group (nrv rows of nx
// nodes, contiguous in the field) is written to the wave's staging buffer st
// (64*L nodes) at the lanes' chunk positions, then read back and stored with
// lane m of each store instruction writing vector m of the group: every
// instruction then writes one contiguous run of 64 vectors (whole 64-byte
// segments), where storing the lanes' own chunks writes every other 16 bytes
// per instruction.  The staging buffer is private to the wave, and a wave's
// LDS accesses complete in issue order, so no barrier is needed.
template <typename T, int L, bool DEV>
__device__ __forceinline__ void x_row_solve_staged(const SweepArgs<T>& a, int d, int rg, int nrv,
                                                   int rw, bool have, int p, int P, int s,
                                                   int valid, const XCoef<T, L>& c, XRow<T, L>& q,
                                                   T* __restrict__ st) {
    using VA = typename Vec16A<T>::type;
    using VU = typename Vec16<T>::type;
    constexpr int W = Vec16<T>::n;
    const int nx = a.nx;
    // the lane's chunk position and count are loop-invariant; laundering them
    // per row group keeps the compiler from hoisting every position test (a
    // scalar mask register pair each) out of the row loop, which otherwise
    // exhausts the scalar registers and spills them into vector lanes
    asm volatile("" : "+v"(s), "+v"(valid));
    if (DEV && have) {  // the chunk's values, read back after the delta solve
        T* dst = st + rw * nx + s;
        if (nx % W == 0) {
#pragma unroll
            for (int v = 0; v < L / W; ++v)
                if (v * W < valid) {
                    VA t;
#pragma unroll
                    for (int w = 0; w < W; ++w) t[w] = q.v[v * W + w];
                    *reinterpret_cast<VA*>(dst + v * W) = t;
                }
        } else {
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (m < valid) dst[m] = q.v[m];
        }
    }
    x_row_compute<T, L, DEV>(nx, have, p, P, s, valid, c, q, st + rw * nx + s);
    if (have) {
        T* dst = st + rw * nx + s;
        if (nx % W == 0) {  // chunk starts on a vector boundary; valid is a multiple of W
#pragma unroll
            for (int v = 0; v < L / W; ++v) {
                if (v * W < valid) {
                    VA t;
#pragma unroll
                    for (int w = 0; w < W; ++w) t[w] = q.v[v * W + w];
                    *reinterpret_cast<VA*>(dst + v * W) = t;
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (m < valid) dst[m] = q.v[m];
        }
    }
    __builtin_amdgcn_wave_barrier();
    const int span = nrv * nx;
    T* gdst = a.u + (long long)d * a.N + (long long)rg * nx;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < L / W; ++i) {
        const int m = (i * 64 + lane) * W;
        if (m + W <= span) {
            const VA t = *reinterpret_cast<const VA*>(st + m);
            VU u;
#pragma unroll
            for (int w = 0; w < W; ++w) u[w] = t[w];
            __builtin_nontemporal_store(u, reinterpret_cast<VU*>(gdst + m));
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (m + w < span) gdst[m + w] = st[m + w];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// Rows [r0, r1) of substrate d in row groups of RW, pipelined: the next
// group's loads are issued before the current group is solved.
template <typename T, int L, bool DEV>
__device__ __forceinline__ void x_rows_pipelined(const SweepArgs<T>& a, int d, int r0, int r1,
                                                 int RW, int rw, bool lane_on, int p, int P, int s,
                                                 int valid, const XCoef<T, L>& c,
                                                 T* __restrict__ st) {
    XRow<T, L> nxt;
    x_row_fetch<T, L>(a, d, r0 + rw, lane_on && r0 + rw < r1, s, valid, c, nxt);
    for (int g = r0; g < r1; g += RW) {
        XRow<T, L> cur = nxt;
        const int r = g + rw;
        if (g + RW < r1) {
            const int rn = r + RW;
            x_row_fetch<T, L>(a, d, rn, lane_on && rn < r1, s, valid, c, nxt);
        }
        const int nrv = r1 - g < RW ? r1 - g : RW;
        x_row_solve_staged<T, L, DEV>(a, d, g, nrv, rw, lane_on && r < r1, p, P, s, valid, c, cur, st);
    }
}

template <typename T, int L, bool DEV, bool FR>
__global__ __launch_bounds__(256) void k_sweep_x_wave(SweepArgs<T> a) {
    const int nx = a.nx;
    const int P = (nx + L - 1) / L;
    const int RW = 64 / P;
    const int lane = threadIdx.x & 63;
    const int rw = lane / P;
    const int p = lane - rw * P;
    const bool lane_on = rw < RW;
    // uniform over the block: keep the tile and substrate in scalar registers
    const int lb = __builtin_amdgcn_readfirstlane(
        xcd_block(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y));
    const int bx = __builtin_amdgcn_readfirstlane(lb % gridDim.x);
    const int d = __builtin_amdgcn_readfirstlane(a.d0 + lb / gridDim.x);
    const int s = p * L;
    const int valid = lane_on ? (nx - s < L ? nx - s : L) : 0;
    const int row0 = (int)a.first;
    const int rows = row0 + (int)a.count;
    const int ngroups = ((int)a.count + RW - 1) / RW;
    const int wstride = gridDim.x * (blockDim.x >> 6);
    XCoef<T, L> c;
    x_coef_setup<T, L, DEV>(a, d, p, P, s, valid, c);
    __shared__ __align__(16) T s_stage[4 * 64 * L];  // one row-group buffer per wave
    T* st = s_stage + (threadIdx.x >> 6) * 64 * L;
    // two-stage pipeline: the loads of the next row group are in flight
    // while the current one is solved
    int g = bx * (blockDim.x >> 6) + (threadIdx.x >> 6);
    XRow<T, L> nxt;
    {
        const int r = row0 + g * RW + rw;
        x_row_fetch<T, L, FR>(a, d, r, g < ngroups && lane_on && r < rows, s, valid, c, nxt);
    }
    for (; g < ngroups; g += wstride) {
        XRow<T, L> cur = nxt;
        const int r = row0 + g * RW + rw;
        const int gn = g + wstride;
        if (gn < ngroups) {
            const int rn = row0 + gn * RW + rw;
            x_row_fetch<T, L, FR>(a, d, rn, lane_on && rn < rows, s, valid, c, nxt);
        }
        const int rg = row0 + g * RW;
        const int nrv = rows - rg < RW ? rows - rg : RW;
        x_row_solve_staged<T, L, DEV>(a, d, rg, nrv, rw, lane_on && r < rows, p, P, s, valid, c, cur, st);
    }
}

// ---- y / z sweeps: one lane per x-column, TX columns per block -------------
// AXIS 1 = y (line stride nx, columns (k,i)), AXIS 2 = z (stride nx*ny,
// columns (j,i)).  Offsets within a substrate are 32-bit (N < 2^31).

// Dirichlet pass after the solve of column (ci, co): the line's end nodes,
// and all its nodes when the line lies on a face parallel to it.  Only the
// threads holding an end node look values up, unless the line is on a face.
template <typename T, int L, int AXIS>
__device__ __forceinline__ void col_dirichlet(T (&x)[L], const SweepArgs<T>& a, int d, int ci,
                                              int co, int s, int n, int p, int P) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const bool lineb = AXIS == 1 ? (ci == 0 || ci == nx - 1 || co == 0 || co == nz - 1)
                                 : (ci == 0 || ci == nx - 1 || co == 0 || co == ny - 1);
    const int s_lo = AXIS == 1 ? dir_slot(ci, 0, co, nx, ny, nz, a.reg)
                               : dir_slot(ci, co, 0, nx, ny, nz, a.reg);
    const int s_hi = AXIS == 1 ? dir_slot(ci, n - 1, co, nx, ny, nz, a.reg)
                               : dir_slot(ci, co, n - 1, nx, ny, nz, a.reg);
    if (lineb) {
        const int s_in = AXIS == 1 ? dir_slot(ci, -1, co, nx, ny, nz, a.reg)
                                   : dir_slot(ci, co, -1, nx, ny, nz, a.reg);
        dirichlet_row<T, L>(x, s, n, true, a.dval[s_lo * a.ns + d], a.dval[s_hi * a.ns + d],
                            a.dval[s_in * a.ns + d]);
    } else {
        if (p == 0) x[0] = a.dval[s_lo * a.ns + d];
        if (p == P - 1) {
            const T v = a.dval[s_hi * a.ns + d];
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (s + m == n - 1) x[m] = v;
        }
    }
}

// Load / store the cnt (<= 0: none) nodes of a thread's chunk at line +
// m*stride; nodes past cnt load as 0.  Full chunks take the unpredicated path.
template <typename T, int L>
__device__ __forceinline__ void load_line_chunk(T (&x)[L], const T* line, size_t stride, int cnt) {
    if (cnt >= L) {
#pragma unroll
        for (int m = 0; m < L; ++m) x[m] = line[(size_t)m * stride];
    } else {
#pragma unroll
        for (int m = 0; m < L; ++m) x[m] = m < cnt ? line[(size_t)m * stride] : T(0);
    }
}

template <typename T, int L>
__device__ __forceinline__ void store_line_chunk(const T (&x)[L], T* line, size_t stride, int cnt) {
    if (cnt >= L) {
#pragma unroll
        for (int m = 0; m < L; ++m) line[(size_t)m * stride] = x[m];
    } else {
#pragma unroll
        for (int m = 0; m < L; ++m)
            if (m < cnt) line[(size_t)m * stride] = x[m];
    }
}

// Where the FP32 column sweep keeps a chunk's values during the increment
// solve: in registers for chunks of 32 or more nodes (their register copy
// still fits, and the LDS copy of a whole block would leave room for fewer
// blocks: at 800^3, y 1.60 against 1.68 ms, z 1.86 against 1.94 ms), in LDS
// below (at 400^3 with 25-node chunks, y 0.371 against 0.389 ms).
__host__ __device__ constexpr bool col_stash_in_regs(int L) { return L >= 32; }

// blockDim.x = TX*P with P = ceil(n/L); grid = (ceil(ncols/TX), ns).
// Dynamic LDS: (2*P*L + TX*P + P) * sizeof(T), plus TX*P*L * sizeof(T) for the
// FP32 form's copy of the values when it is kept in LDS.
template <typename T, int L, int AXIS, bool DEV, int TX>
__global__ __launch_bounds__(1024) void k_sweep_col(SweepArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int n = AXIS == 1 ? ny : nz;
    const int P = (n + L - 1) / L;
    T* s_inv = reinterpret_cast<T*>(smem);
    T* s_cp = s_inv + P * L;
    T* s_z = s_cp + P * L;
    T* s_g = s_z + (size_t)TX * P;

    const int t = threadIdx.x;
    const int lane = t % TX;
    const int p = t / TX;
    // uniform over the block: keep the tile and substrate in scalar registers
    const int lb = __builtin_amdgcn_readfirstlane(
        xcd_runs(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y, a.runs));
    const int d = __builtin_amdgcn_readfirstlane(a.d0 + lb / gridDim.x);
    stage_coefs<T>(s_inv, s_cp, a.inv + (long long)d * n, a.cp + (long long)d * n, n, P * L, t,
                   blockDim.x);
    const int c = (int)a.first + __builtin_amdgcn_readfirstlane(lb % gridDim.x) * TX + lane;
    const bool act = c < (int)(a.first + a.count);
    int ci = 0, co = 0;  // i and the other transverse coordinate (k for y, j for z)
    if (act) {
        co = c / nx;
        ci = c - co * nx;
    }
    T* ub = a.u + (long long)d * a.N;
    const unsigned base = AXIS == 1 ? (unsigned)(co * ny * nx + ci) : (unsigned)c;
    const size_t stride = AXIS == 1 ? (size_t)nx : (size_t)nx * ny;
    const int s = p * L;
    // this thread's chunk: node m at line + m*stride (a uniform stride, so
    // each address is one 64-bit add)
    T* line = ub + (size_t)base + (size_t)s * stride;
    const int cnt = act ? n - s : 0;
    T x[L];
    load_line_chunk<T, L>(x, line, stride, cnt);
    if (DEV) {
        // delta form: the chunk's outer neighbours come from the field (the
        // neighbouring chunks' threads load them too, so they hit in cache);
        // the values wait for x = v + delta in LDS, or with chunks of 32 in
        // registers (col_stash_in_regs)
        const T vl = (act && s > 0) ? *(line - stride) : T(0);
        const T vr = (act && s + L < n) ? line[(size_t)L * stride] : T(0);
        T* s_v = s_g + P;
        T vsave[col_stash_in_regs(L) ? L : 1];
        if constexpr (col_stash_in_regs(L)) {
#pragma unroll
            for (int m = 0; m < L; ++m) vsave[m] = x[m];
        } else {
            col_stash<T, L>(s_v, x, p, lane, TX);
        }
        delta_residual<T, L>(x, vl, vr, s, act ? n : 0, a.k[d], a.k[a.ns + d], a.beta[d]);
        __syncthreads();  // coefficient staging complete
        chunk_solve<T, L>(x, act, p, P, lane, TX, s_inv, s_cp, a.k[d], s_z, s_g);
        if constexpr (col_stash_in_regs(L)) {
#pragma unroll
            for (int m = 0; m < L; ++m) x[m] = vsave[m] + x[m];
        } else {
            col_unstash_add<T, L>(s_v, x, p, lane, TX);
        }
    } else {
        __syncthreads();  // coefficient staging complete
        chunk_solve<T, L>(x, act, p, P, lane, TX, s_inv, s_cp, a.k[d], s_z, s_g);
    }
    if (act && ((a.active >> d) & 1ull)) col_dirichlet<T, L, AXIS>(x, a, d, ci, co, s, n, p, P);
    store_line_chunk<T, L>(x, line, stride, cnt);
}

}  // namespace bfa
