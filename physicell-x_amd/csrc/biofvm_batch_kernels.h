// biofvm_batch_kernels.h -- gfx950 device code for batches of independent
// microenvironments of different shapes (the ensemble and serving forms).
//
// Every member has its own device field, coefficients and Dirichlet table,
// described by an EnvDesc.  A sweep over the whole batch is one persistent
// kernel: its blocks (x sweep: its wavefronts) take work items from an atomic
// ticket counter until the item list is exhausted, so members of very
// different sizes are load-balanced dynamically and every wave reaches the
// exit once the counter passes the item count.  Items are listed largest
// first.  The per-line arithmetic is the single-environment one
// (biofvm_kernels.h), so a member's result does not depend on what else is in
// the batch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "biofvm_kernels.h"

namespace bfa {

template <typename T>
struct EnvDesc {
    T* u;              // field, device layout ((d*nz + k)*ny + j)*nx + i
    const T* inv[3];   // per axis [ns][n] 1/den
    const T* cp[3];    // per axis [ns][n] c'
    const T* k;        // [ns]
    const T* beta;     // [ns]
    const T* dval;     // [7][ns]
    long long N;
    int nx, ny, nz, ns;
    unsigned reg;
    unsigned long long active;
    const double* params;  // [5][ns]: D, decay, b, s, 1/s (FP32 deviation basis)
    int steps;  // steps of the current run this member takes; items of later steps are skipped
    int pad;
};

// x sweep item: rows [row0, row0 + nrows) of substrate d of member env.
struct XItem {
    int env, d, row0, nrows;
};
// y / z sweep item: columns [c0, c0 + TX) of substrate d of member env.
struct ColItem {
    int env, d, c0, pad;
};

template <typename T>
__device__ __forceinline__ SweepArgs<T> args_of(const EnvDesc<T>& e, int axis) {
    SweepArgs<T> a;
    a.u = e.u;
    a.inv = e.inv[axis];
    a.cp = e.cp[axis];
    a.k = e.k;
    a.beta = e.beta;
    a.dval = e.dval;
    a.active = e.active;
    a.N = e.N;
    a.nx = e.nx;
    a.ny = e.ny;
    a.nz = e.nz;
    a.ns = e.ns;
    a.reg = e.reg;
    a.d0 = 0;
    a.first = 0;
    a.count = 0;
    return a;
}

// Persistent batched x sweep.  Each wavefront takes XItems from ticket[0]
// until *nitems is reached.  Members need nx <= 64*L.
template <typename T, int L, bool DEV>
__global__ __launch_bounds__(256) void kb_sweep_x(const EnvDesc<T>* __restrict__ envs,
                                                  const XItem* __restrict__ items,
                                                  const int* __restrict__ nitems,
                                                  int* __restrict__ ticket, int step) {
    const int lane = threadIdx.x & 63;
    const int n_items = *nitems;
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(ticket, 1);
        it = __shfl(it, 0, 64);
        if (it >= n_items) break;
        const XItem w = items[it];
        if (step >= envs[w.env].steps) continue;
        const SweepArgs<T> a = args_of<T>(envs[w.env], 0);
        const int nx = a.nx;
        const int P = (nx + L - 1) / L;
        const int RW = 64 / P;
        if (RW == 0) continue;  // row longer than 64 chunks: the host never lists such items
        const int rw = lane / P;
        const int p = lane - rw * P;
        const bool lane_on = rw < RW;
        const int s = p * L;
        const int valid = lane_on ? (nx - s < L ? nx - s : L) : 0;
        T inv[L], cp[L];
#pragma unroll
        for (int m = 0; m < L; ++m) {
            inv[m] = m < valid ? a.inv[w.d * nx + s + m] : T(0);
            cp[m] = m < valid ? a.cp[w.d * nx + s + m] : T(0);
        }
        const bool vec = ((nx * (int)sizeof(T)) & 15) == 0 && ((L * (int)sizeof(T)) & 15) == 0;
        const int rend = w.row0 + w.nrows;
        for (int r0 = w.row0; r0 < rend; r0 += RW) {
            const int r = r0 + rw;
            const bool have = lane_on && r < rend;
            if (vec)
                x_row_group<T, L, DEV, true>(a, w.d, r, have, p, P, s, valid, inv, cp);
            else
                x_row_group<T, L, DEV, false>(a, w.d, r, have, p, P, s, valid, inv, cp);
        }
    }
}

// Persistent batched y (AXIS 1) or z (AXIS 2) sweep.  Each block takes
// ColItems from ticket[0]; blockDim.x = TX * PMAX with PMAX >= ceil(n/L) for
// every member.  Dynamic LDS: (2*PMAX*L + TX*PMAX + PMAX) * sizeof(T) + 16.
template <typename T, int L, int AXIS, bool DEV, int TX>
__global__ __launch_bounds__(1024) void kb_sweep_col(const EnvDesc<T>* __restrict__ envs,
                                                     const ColItem* __restrict__ items,
                                                     const int* __restrict__ nitems,
                                                     int* __restrict__ ticket, int PMAX,
                                                     int step) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* s_inv = reinterpret_cast<T*>(smem);
    T* s_cp = s_inv + PMAX * L;
    T* s_z = s_cp + PMAX * L;
    T* s_g = s_z + (size_t)TX * PMAX;
    int* s_item = reinterpret_cast<int*>(s_g + PMAX);
    const int t = threadIdx.x;
    const int lane = t % TX;
    const int p = t / TX;
    const int n_items = *nitems;
    for (;;) {
        __syncthreads();  // previous item's LDS readers are done
        if (t == 0) *s_item = atomicAdd(ticket, 1);
        __syncthreads();
        const int it = *s_item;
        if (it >= n_items) break;
        const ColItem w = items[it];
        if (step >= envs[w.env].steps) continue;  // uniform over the block
        const SweepArgs<T> a = args_of<T>(envs[w.env], AXIS);
        const int nx = a.nx, ny = a.ny, nz = a.nz;
        const int n = AXIS == 1 ? ny : nz;
        const int P = (n + L - 1) / L;
        const int d = w.d;
        stage_coefs<T>(s_inv, s_cp, a.inv + (long long)d * n, a.cp + (long long)d * n, n, P * L, t,
                       blockDim.x);
        const long long ncols = AXIS == 1 ? (long long)nz * nx : (long long)ny * nx;
        const long long c = (long long)w.c0 + lane;
        const bool act = p < P && c < ncols;
        long long base = 0, stride;
        int ci = 0, co = 0;
        if (act) {
            co = (int)(c / nx);
            ci = (int)(c - (long long)co * nx);
        }
        if (AXIS == 1) {
            base = (long long)d * a.N + (long long)co * ny * nx + ci;
            stride = nx;
        } else {
            base = (long long)d * a.N + c;
            stride = (long long)nx * ny;
        }
        const int s = p * L;
        T x[L];
#pragma unroll
        for (int m = 0; m < L; ++m) x[m] = (act && s + m < n) ? a.u[base + (s + m) * stride] : T(0);
        if (DEV) {
            const T bt = a.beta[d];
#pragma unroll
            for (int m = 0; m < L; ++m) x[m] = (s + m < n) ? x[m] - bt : T(0);
        }
        __syncthreads();  // staging done, s_item read by all
        chunk_solve<T, L>(x, act, p, P, lane, TX, s_inv, s_cp, a.k[d], s_z, s_g);
        const bool dact = (a.active >> d) & 1ull;
        if (act && dact) {
            const bool lineb = AXIS == 1 ? (ci == 0 || ci == nx - 1 || co == 0 || co == nz - 1)
                                         : (ci == 0 || ci == nx - 1 || co == 0 || co == ny - 1);
            const int s_lo = AXIS == 1 ? dir_slot(ci, 0, co, nx, ny, nz, a.reg)
                                       : dir_slot(ci, co, 0, nx, ny, nz, a.reg);
            const int s_hi = AXIS == 1 ? dir_slot(ci, n - 1, co, nx, ny, nz, a.reg)
                                       : dir_slot(ci, co, n - 1, nx, ny, nz, a.reg);
            const int s_in = AXIS == 1 ? dir_slot(ci, -1, co, nx, ny, nz, a.reg)
                                       : dir_slot(ci, co, -1, nx, ny, nz, a.reg);
            dirichlet_row<T, L>(x, s, n, lineb, a.dval[s_lo * a.ns + d], a.dval[s_hi * a.ns + d],
                                a.dval[s_in * a.ns + d]);
        }
        if (act) {
#pragma unroll
            for (int m = 0; m < L; ++m)
                if (s + m < n) a.u[base + (s + m) * stride] = x[m];
        }
    }
}

// Agent entries over the whole batch: entry q updates the value at absolute
// device address u_of[q] with its agents' maps a0..a1 (ascending agent
// order): rho = (rho + c1)/c2 + e2 in FP64; FP32 members lift v to
// u = b + s*v in FP64 for the update and store (u - b)/s.
struct AgentEntry {
    unsigned long long addr;  // element address
    int a0, a1;
    int env, d;
};

template <typename T>
__global__ void kb_agents_f64(const EnvDesc<T>* __restrict__ envs,
                              const AgentEntry* __restrict__ ent, const int* __restrict__ nent,
                              const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ e2, int step) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *nent) return;
    const AgentEntry en = ent[q];
    if (step >= envs[en.env].steps) return;
    double* p = reinterpret_cast<double*>(en.addr);
    double rho = *p;
    for (int a = en.a0; a < en.a1; ++a) {
        rho += c1[a];
        rho /= c2[a];
        rho += e2[a];
    }
    *p = rho;
}

template <typename T>
__global__ void kb_agents_f32(const EnvDesc<T>* __restrict__ envs,
                              const AgentEntry* __restrict__ ent, const int* __restrict__ nent,
                              const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ e2, int step) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *nent) return;
    const AgentEntry en = ent[q];
    const EnvDesc<T>& e = envs[en.env];
    if (step >= e.steps) return;
    const double b = e.params[2 * e.ns + en.d], s = e.params[3 * e.ns + en.d],
                 rs = e.params[4 * e.ns + en.d];
    float* p = reinterpret_cast<float*>(en.addr);
    double rho = fma((double)*p, s, b);
    for (int a = en.a0; a < en.a1; ++a) {
        rho += c1[a];
        rho /= c2[a];
        rho += e2[a];
    }
    *p = (float)((rho - b) * rs);
}

// Zero `n` ints (graph-safe replacement for a memset node).
__global__ void kb_zero(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace bfa
