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
    double W;   // voxel volume dx^3 (internalised-substrate tracking)
};

// x sweep item: rows [row0, row0 + nrows) of substrate d of member env.
struct XItem {
    int env, d, row0, nrows;
};
// y / z sweep item: columns [c0, c0 + ntiles*W) of substrate d of member
// env, solved as ntiles tiles of W = col_tile_width(block, P) columns.
struct ColItem {
    int env, d, c0, ntiles;
};

// Columns per tile for lines of P chunks in a block of T threads: the
// widest multiple of 16 with width * P <= T (T >= 16 * P is required).
// Short lines get wide tiles, so every item fills the block.
__host__ __device__ inline int col_tile_width(int T, int P) {
    const int w = (T / P) & ~15;
    return w < 16 ? 16 : w;
}

template <typename T>
__device__ __forceinline__ SweepArgs<T> args_of(const EnvDesc<T>& e, int axis) {
    SweepArgs<T> a;
    a.u = e.u;
    a.inv = e.inv[axis];
    a.cp = e.cp[axis];
    a.k = e.k;
    a.beta = e.beta;
    a.dval = e.dval;
    a.plane = nullptr;  // batch members keep the face table
    a.shell = 0;
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
    a.runs = 1;
    return a;
}

// Persistent batched x sweep.  Each wavefront takes XItems from ticket[0]
// until *nitems is reached; the ticket of its next item is drawn while the
// current one runs, and within an item the loads of the next row group are in
// flight while the current one is solved.  Members need nx <= 64*L.
template <typename T, int L, bool DEV>
__global__ __launch_bounds__(256) void kb_sweep_x(const EnvDesc<T>* __restrict__ envs,
                                                  const XItem* __restrict__ items,
                                                  const int* __restrict__ nitems,
                                                  int* __restrict__ ticket, int step) {
    const int lane = threadIdx.x & 63;
    const int n_items = *nitems;
    __shared__ __align__(16) T s_stage[4 * 64 * L];  // one row-group buffer per wave
    T* st = s_stage + (threadIdx.x >> 6) * 64 * L;
    int nxt = 0;
    if (lane == 0) nxt = atomicAdd(ticket, 1);
    for (;;) {
        const int it = __builtin_amdgcn_readfirstlane(nxt);  // lane 0's ticket
        if (it >= n_items) break;
        if (lane == 0) nxt = atomicAdd(ticket, 1);
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
        XCoef<T, L> c;
        x_coef_setup<T, L, DEV>(a, w.d, p, P, s, valid, c);
        x_rows_pipelined<T, L, DEV>(a, w.d, w.row0, w.row0 + w.nrows, RW, rw, lane_on, p, P, s, valid, c, st);
    }
}

// Persistent batched y (AXIS 1) or z (AXIS 2) sweep over every member.  Each
// block of T threads takes ColItems from ticket[0]; an item whose lines have
// P chunks runs as tiles of W = col_tile_width(T, P) columns x P chunk rows
// (T >= 16 * PMAX).  The ticket of the next item is drawn while the current
// one runs (two LDS slots, alternating), the coefficients are staged once per
// item, and the loads of tile q+1 are issued before tile q is solved.
// Dynamic LDS: (2*PMAX*L + T + PMAX + 8) * sizeof(T) + 16, plus T*L*sizeof(T)
// for the FP32 form's copy of the values (batch_col_lds).  Chunks of 128 bytes
// or more take blocks of at most kb_col_max_threads(L * sizeof(T)) = 512
// threads, which leaves them the registers to stay out of scratch.
__host__ __device__ constexpr int kb_col_max_threads(int chunk_bytes) {
    return chunk_bytes >= 128 ? 512 : 1024;
}

// Column Dirichlet (D2 / D3) of the batch column sweeps: the node values
// col_dirichlet writes, read from the substrate's seven slot values staged in
// LDS (s_dv) with the face selection of dir_slot.  A line lying in a face
// (lineb) is a Dirichlet line: every node takes the value of its own slot,
// the two end nodes those of their slots; other lines only have their end
// nodes on a face.  Nodes past the line end are overwritten too (they are
// not stored), which leaves one position test per node instead of three.
template <typename T, int L, int AXIS>
__device__ __forceinline__ void col_dirichlet_staged(T (&x)[L], const T* s_dv, unsigned reg, int nx,
                                                     int ny, int nz, int ci, int co, int s, int n,
                                                     int p, int P) {
    const bool lineb = AXIS == 1 ? (ci == 0 || ci == nx - 1 || co == 0 || co == nz - 1)
                                 : (ci == 0 || ci == nx - 1 || co == 0 || co == ny - 1);
    const int lo = AXIS == 1 ? dir_slot(ci, 0, co, nx, ny, nz, reg) : dir_slot(ci, co, 0, nx, ny, nz, reg);
    const int hi = AXIS == 1 ? dir_slot(ci, n - 1, co, nx, ny, nz, reg) : dir_slot(ci, co, n - 1, nx, ny, nz, reg);
    if (lineb) {
        const T vin = s_dv[AXIS == 1 ? dir_slot(ci, -1, co, nx, ny, nz, reg) : dir_slot(ci, co, -1, nx, ny, nz, reg)];
#pragma unroll
        for (int m = 0; m < L; ++m) x[m] = vin;
    }
    if (p == 0) x[0] = s_dv[lo];
    if (p == P - 1) {
        const T v = s_dv[hi];
#pragma unroll
        for (int m = 0; m < L; ++m)
            if (s + m == n - 1) x[m] = v;
    }
}

template <typename T, int L, int AXIS, bool DEV>
__global__ __launch_bounds__(kb_col_max_threads(L * sizeof(T))) void kb_sweep_col(const EnvDesc<T>* __restrict__ envs,
                                                     const ColItem* __restrict__ items,
                                                     const int* __restrict__ nitems,
                                                     int* __restrict__ ticket, int PMAX,
                                                     int step) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* s_inv = reinterpret_cast<T*>(smem);
    T* s_cp = s_inv + PMAX * L;
    T* s_z = s_cp + PMAX * L;
    T* s_g = s_z + blockDim.x;
    int* s_item = reinterpret_cast<int*>(s_g + PMAX);
    T* s_dv = reinterpret_cast<T*>(s_item + 4);  // the item substrate's 7 Dirichlet slot values
    const int t = threadIdx.x;
    const int n_items = *nitems;
    if (t == 0) s_item[0] = atomicAdd(ticket, 1);
    for (int slot = 0;; slot ^= 1) {
        __syncthreads();  // s_item[slot] written; the previous item's LDS readers are done
        const int it = __builtin_amdgcn_readfirstlane(s_item[slot]);
        if (it >= n_items) break;
        int nxt = 0;
        if (t == 0) nxt = atomicAdd(ticket, 1);
        const ColItem w = items[it];
        if (step >= envs[w.env].steps) {  // uniform over the block
            if (t == 0) s_item[slot ^ 1] = nxt;
            continue;
        }
        const SweepArgs<T> a = args_of<T>(envs[w.env], AXIS);
        const int nx = a.nx, ny = a.ny, nz = a.nz;
        int n = AXIS == 1 ? ny : nz;
        int P = (n + L - 1) / L;
        int W = col_tile_width(blockDim.x, P);
        int lane = t % W;
        int p = t / W;
        const int d = w.d;
        stage_coefs<T>(s_inv, s_cp, a.inv + (long long)d * n, a.cp + (long long)d * n, n, P * L, t,
                       blockDim.x);
        if (t < 7) s_dv[t] = a.dval[t * a.ns + d];
        const int ncols = AXIS == 1 ? nz * nx : ny * nx;
        const size_t stride = AXIS == 1 ? (size_t)nx : (size_t)nx * ny;
        T* ub = a.u + (long long)d * a.N;
        int s = p * L;
        const T bt = DEV ? a.beta[d] : T(0);
        const T kd = a.k[d];
        const T lam = DEV ? a.k[a.ns + d] : T(0);
        const bool dact = (a.active >> d) & 1ull;
        // column c of the item: its (ci, co), offset of its first node, and
        // whether this thread holds part of it
        auto col_of = [&](int c, int& ci, int& co, unsigned& base) -> bool {
            const bool on = p < P && c < ncols;
            co = on ? c / nx : 0;
            ci = on ? c - co * nx : 0;
            base = AXIS == 1 ? (unsigned)(co * ny * nx + ci) : (unsigned)c;
            return on;
        };
        T xn[L];
        T hl_n = T(0), hr_n = T(0);  // FP32 delta form: the chunk's outer neighbours
        int ci_n, co_n;
        unsigned base_n;
        bool act_n = col_of(w.c0 + lane, ci_n, co_n, base_n);
        auto fetch = [&](bool on, unsigned bse) {
            const T* line = ub + (size_t)bse + (size_t)s * stride;
            load_line_chunk<T, L>(xn, line, stride, on ? n - s : 0);
            if (DEV) {
                hl_n = (on && s > 0) ? *(line - stride) : T(0);
                hr_n = (on && s + L < n) ? line[(size_t)L * stride] : T(0);
            }
        };
        fetch(act_n, base_n);
        if (t == 0) s_item[slot ^ 1] = nxt;
        for (int q = 0; q < w.ntiles; ++q) {
            // the thread's place in the tile (chunk position, line length,
            // chunk and column index, chunk count, tile width) is invariant
            // over the item's tiles; laundering it per tile keeps the tests
            // on it (a scalar mask register pair each) from being hoisted out
            // of the tile loop and spilling the scalar registers
            asm volatile("" : "+v"(s), "+v"(n), "+v"(p), "+v"(P), "+v"(lane), "+v"(W));
            T x[L];
#pragma unroll
            for (int m = 0; m < L; ++m) x[m] = xn[m];
            const T hl = hl_n, hr = hr_n;
            const unsigned base = base_n;
            const int ci = ci_n, co = co_n;
            const bool act = act_n;
            if (q + 1 < w.ntiles) {
                act_n = col_of(w.c0 + (q + 1) * W + lane, ci_n, co_n, base_n);
                fetch(act_n, base_n);
            }
            if (DEV) {
                // the values wait in registers for x = v + delta (no extra VGPRs at
                // L <= 32: the chunk's values are live across the solve anyway)
                T vsave[L];
#pragma unroll
                for (int m = 0; m < L; ++m) vsave[m] = x[m];
                delta_residual<T, L>(x, hl, hr, s, act ? n : 0, kd, lam, bt);
                __syncthreads();  // staging done / the previous tile's fold readers are done
                chunk_solve<T, L>(x, act, p, P, lane, W, s_inv, s_cp, kd, s_z, s_g);
#pragma unroll
                for (int m = 0; m < L; ++m) x[m] = vsave[m] + x[m];
            } else {
                __syncthreads();  // staging done / the previous tile's fold readers are done
                chunk_solve<T, L>(x, act, p, P, lane, W, s_inv, s_cp, kd, s_z, s_g);
            }
            if (act && dact) col_dirichlet_staged<T, L, AXIS>(x, s_dv, a.reg, nx, ny, nz, ci, co, s, n, p, P);
            store_line_chunk<T, L>(x, ub + (size_t)base + (size_t)s * stride, stride, act ? n - s : 0);
        }
    }
}

// ---- fused x + y sweeps of whole k-planes held in LDS ------------------------
// A k-plane of a member (nx*ny values, contiguous in the device layout) is
// read from HBM once into LDS, swept along x (Dirichlet D0, the row solves,
// D1) and along y (the column solves, D2) there, and written back once, in
// place of the two HBM round trips of separate x and y sweeps.  Used for the
// members whose plane fits the block's plane buffer; the others take the
// separate sweeps.  Per-line arithmetic is that of the separate sweeps
// (x_row_compute, chunk_solve, col_dirichlet) with x chunks of LX nodes and
// y chunks of LY nodes.

// xy item: planes [k0, k0 + nk) of substrate d of member env.
struct XYItem {
    int env, d, k0, nk;
};

// Threads per block and 16-byte vectors per thread of the plane buffer
// (the next plane's prefetch registers): planes of up to threads * vecs * 16
// bytes (128 KiB) are fused.
template <typename T>
struct XYShape {
    static constexpr int threads = sizeof(T) == 4 ? 512 : 256;
    static constexpr int vecs = sizeof(T) == 4 ? 16 : 32;
    static constexpr int max_bytes = threads * vecs * 16;
};

// LX is odd, so the lanes of a row (LX nodes apart) read distinct LDS banks.
// Dynamic LDS: (PLANE + LX) + 2*PYMAX*LY + threads + PYMAX values of T, + 16.
template <typename T, int LX, int LY, bool DEV>
__global__ __launch_bounds__(XYShape<T>::threads) void kb_sweep_xy(
    const EnvDesc<T>* __restrict__ envs, const XYItem* __restrict__ items,
    const int* __restrict__ nitems, int* __restrict__ ticket, int PLANE, int PYMAX, int step) {
    using V = typename Vec16<T>::type;  // unaligned 16-byte vector (global side)
    typedef T VA __attribute__((ext_vector_type(16 / sizeof(T))));  // aligned (LDS side)
    constexpr int VW = 16 / sizeof(T);
    constexpr int NV = XYShape<T>::vecs;
    extern __shared__ __align__(16) unsigned char smem[];
    T* plane = reinterpret_cast<T*>(smem);
    T* s_inv = plane + ((PLANE + LX + VW - 1) / VW) * VW;
    T* s_cp = s_inv + PYMAX * LY;
    T* s_z = s_cp + PYMAX * LY;
    T* s_g = s_z + blockDim.x;
    int* s_item = reinterpret_cast<int*>(s_g + PYMAX);
    const int t = threadIdx.x;
    const int nthr = blockDim.x;
    const int wave = t >> 6, lane64 = t & 63, nwaves = nthr >> 6;
    const int n_items = *nitems;
    if (t == 0) s_item[0] = atomicAdd(ticket, 1);
    for (int slot = 0;; slot ^= 1) {
        __syncthreads();  // s_item[slot] written; the previous item's LDS readers are done
        const int it = __builtin_amdgcn_readfirstlane(s_item[slot]);
        if (it >= n_items) break;
        int nxt = 0;
        if (t == 0) nxt = atomicAdd(ticket, 1);
        const XYItem w = items[it];
        if (step >= envs[w.env].steps) {  // uniform over the block
            if (t == 0) s_item[slot ^ 1] = nxt;
            continue;
        }
        const SweepArgs<T> ax = args_of<T>(envs[w.env], 0);
        const SweepArgs<T> ay = args_of<T>(envs[w.env], 1);
        const int nx = ax.nx, ny = ax.ny, d = w.d;
        const int nxy = nx * ny;
        const int nvec = (nxy + VW - 1) / VW;
        // x: lane p of a row owns nodes [p*LX, p*LX + LX); RW rows per wave
        const int PX = (nx + LX - 1) / LX;
        const int RW = 64 / PX;
        const int rw = lane64 / PX;
        const int px = lane64 - rw * PX;
        const bool xlane = rw < RW;
        const int sx = px * LX;
        const int xvalid = xlane ? (nx - sx < LX ? nx - sx : LX) : 0;
        // y: tiles of W columns x PY chunk rows
        const int PY = (ny + LY - 1) / LY;
        const int W = col_tile_width(nthr, PY);
        const int ly = t % W, py = t / W;
        const int sy = py * LY;
        stage_coefs<T>(s_inv, s_cp, ay.inv + (long long)d * ny, ay.cp + (long long)d * ny, ny, PY * LY,
                       t, nthr);
        const T bt = DEV ? ay.beta[d] : T(0);
        const T ky = ay.k[d];
        const T lam = DEV ? ay.k[ay.ns + d] : T(0);
        const bool dact = (ay.active >> d) & 1ull;
        if (t == 0) s_item[slot ^ 1] = nxt;
        // the plane's 16-byte vectors v = t + q*nthr, q < NV; the last one may
        // run past the plane end (into the next plane or the field slack),
        // which lands in the plane buffer's padding
        V pre[NV];
        const V* src = reinterpret_cast<const V*>(ax.u + (long long)d * ax.N + (long long)w.k0 * nxy);
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int v = t + q * nthr;
            if (v < nvec) pre[q] = __builtin_nontemporal_load(src + v);
        }
        for (int k = w.k0; k < w.k0 + w.nk; ++k) {
            __syncthreads();  // the previous plane's store phase has read the buffer
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const int v = t + q * nthr;
                if (v < nvec) {
                    VA o;
#pragma unroll
                    for (int e = 0; e < VW; ++e) o[e] = pre[q][e];
                    reinterpret_cast<VA*>(plane)[v] = o;
                }
            }
            if (k + 1 < w.k0 + w.nk) {  // in flight while this plane is solved
                const V* nsrc = reinterpret_cast<const V*>(ax.u + (long long)d * ax.N + (long long)(k + 1) * nxy);
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const int v = t + q * nthr;
                    if (v < nvec) pre[q] = __builtin_nontemporal_load(nsrc + v);
                }
            }
            __syncthreads();
            // x solves of the plane's rows (row r of the sweep numbering is k*ny + j);
            // the lane's coefficient state is rebuilt per plane so that it is
            // not held through the y solves
            XCoef<T, LX> cx;
            x_coef_setup<T, LX, DEV>(ax, d, px, PX, sx, xvalid, cx);
            for (int j0 = wave * RW; j0 < ny; j0 += nwaves * RW) {
                const int j = j0 + rw;
                const bool have = xlane && j < ny;
                XRow<T, LX> q;
                // nodes past the row end read as 0 (past the plane's last row
                // the buffer holds stale LDS contents, which may be NaN)
#pragma unroll
                for (int m = 0; m < LX; ++m) q.v[m] = (have && m < xvalid) ? plane[j * nx + sx + m] : T(0);
                x_row_dir_fetch<T, LX>(ax, d, k * ny + j, have, cx, q);
                x_row_compute<T, LX, DEV>(nx, have, px, PX, sx, xvalid, cx, q, plane + j * nx + sx);
                if (have) {
#pragma unroll
                    for (int m = 0; m < LX; ++m)
                        if (m < xvalid) plane[j * nx + sx + m] = q.v[m];
                }
            }
            __syncthreads();  // the plane's rows are solved
            // y solves of the plane's columns, W at a time
            for (int c0 = 0; c0 < nx; c0 += W) {
                const int i = c0 + ly;
                const bool act = py < PY && i < nx;
                T x[LY];
                const T* col = plane + sy * nx + i;
                if (DEV) {
                    // the plane keeps v for x = v + delta: the column's own
                    // nodes are rewritten only by this thread, after the read
#pragma unroll
                    for (int m = 0; m < LY; ++m) x[m] = (act && sy + m < ny) ? col[m * nx] : T(0);
                    const T hl = (act && sy > 0) ? col[-nx] : T(0);
                    const T hr = (act && sy + LY < ny) ? col[LY * nx] : T(0);
                    delta_residual<T, LY>(x, hl, hr, sy, act ? ny : 0, ky, lam, bt);
                    __syncthreads();  // the previous tile's fold readers are done
                    chunk_solve<T, LY>(x, act, py, PY, ly, W, s_inv, s_cp, ky, s_z, s_g);
#pragma unroll
                    for (int m = 0; m < LY; ++m)
                        if (act && sy + m < ny) x[m] = col[m * nx] + x[m];
                } else {
#pragma unroll
                    for (int m = 0; m < LY; ++m) x[m] = (act && sy + m < ny) ? col[m * nx] : T(0);
                    __syncthreads();  // the previous tile's fold readers are done
                    chunk_solve<T, LY>(x, act, py, PY, ly, W, s_inv, s_cp, ky, s_z, s_g);
                }
                if (act && dact) col_dirichlet<T, LY, 1>(x, ay, d, i, k, sy, ny, py, PY);
                if (act) {
                    T* cw = plane + sy * nx + i;
#pragma unroll
                    for (int m = 0; m < LY; ++m)
                        if (sy + m < ny) cw[m * nx] = x[m];
                }
            }
            __syncthreads();
            // store the plane: whole vectors, the plane's last partial one by element
            T* dst = ax.u + (long long)d * ax.N + (long long)k * nxy;
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const int v = t + q * nthr;
                if (v < nvec) {
                    const VA o = reinterpret_cast<const VA*>(plane)[v];
                    if ((v + 1) * VW <= nxy) {
                        V u;
#pragma unroll
                        for (int e = 0; e < VW; ++e) u[e] = o[e];
                        __builtin_nontemporal_store(u, reinterpret_cast<V*>(dst) + v);
                    } else {
#pragma unroll
                        for (int e = 0; e < VW; ++e)
                            if (v * VW + e < nxy) dst[v * VW + e] = o[e];
                    }
                }
            }
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

// Agent step of the batch: one lane per (segment, substrate) entry walks its
// agents in order, as k_agents.  trk[a] (agent-entry order) is the address of
// the agent's internalised total for the entry's substrate, or 0 when its
// member does not track; e1[a] is the agent's net export for that substrate.
template <typename T>
__global__ void kb_agents_f64(const EnvDesc<T>* __restrict__ envs,
                              const AgentEntry* __restrict__ ent, const int* __restrict__ nent,
                              const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ e2, const unsigned long long* __restrict__ trk,
                              const double* __restrict__ e1, int step) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *nent) return;
    const AgentEntry en = ent[q];
    const EnvDesc<T>& e = envs[en.env];
    if (step >= e.steps) return;
    double* p = reinterpret_cast<double*>(en.addr);
    double rho = *p;
    for (int a = en.a0; a < en.a1; ++a) {
        double* I = reinterpret_cast<double*>(trk[a]);
        if (I) track_uptake(I, rho, c1[a], c2[a], e.W);
        rho += c1[a];
        rho /= c2[a];
        rho += e2[a];
        if (I) *I = *I - e1[a];
    }
    *p = rho;
}

template <typename T>
__global__ void kb_agents_f32(const EnvDesc<T>* __restrict__ envs,
                              const AgentEntry* __restrict__ ent, const int* __restrict__ nent,
                              const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ e2, const unsigned long long* __restrict__ trk,
                              const double* __restrict__ e1, int step) {
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
        double* I = reinterpret_cast<double*>(trk[a]);
        if (I) track_uptake(I, rho, c1[a], c2[a], e.W);
        rho += c1[a];
        rho /= c2[a];
        rho += e2[a];
        if (I) *I = *I - e1[a];
    }
    *p = (float)((rho - b) * rs);
}

// Zero `n` ints (graph-safe replacement for a memset node).
__global__ void kb_zero(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace bfa
