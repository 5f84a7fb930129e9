// biofvm_amd.hip -- host side of the C ABI in include/biofvm_amd.h.
//
// One bfa_env owns one microenvironment resident on the current HIP device:
// the SoA field, the per-axis Thomas coefficients (rebuilt on the device in
// FP64 whenever dt, the coefficients or the FP32 basis change), the Dirichlet
// face table, the agents sorted into per-voxel segments, a HIP stream and a
// captured one-step HIP graph.  See biofvm_kernels.h for the sweep design.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "biofvm_amd.h"
#include "biofvm_aux_kernels.h"
#include "biofvm_batch_kernels.h"
#include "biofvm_kernels.h"
#include "biofvm_transfer.inc"

namespace {

thread_local std::string g_create_error;

constexpr int kMaxEdge = 1024;
constexpr int kMaxSubstrates = 64;
constexpr int kLs[] = {4, 8, 16, 32, 64};
// column sweeps also take 25-node chunks (16 chunks of a 400-node line)
constexpr int kColLs[] = {4, 8, 16, 25, 32, 64};
// the wave-local x sweep also takes 20 and 28 (whole 16-byte vectors in FP32
// and FP64): 20 chunks of a 400-node row put three rows in a wavefront
// (60 of 64 lanes, against one row of 50 lanes with 8-node chunks), 29 chunks
// of an 800-node row two (58 lanes)
constexpr int kXWaveLs[] = {4, 8, 16, 20, 28, 32};

}  // namespace

// Elements of zeroed slack after every device field: the x sweeps load whole
// row chunks (up to 31 nodes past a row's end) and never store there.
constexpr size_t kFieldSlack = 64;

struct bfa_env {
    int nx = 0, ny = 0, nz = 0, ns = 0;
    double dx = 0.0;
    int prec = BFA_FP64;
    long long N = 0;
    size_t tsize = 8;
    int device = 0;
    int max_lds = 65536;
    int cus = 256;  // compute units of the device
    int xw_occ = 1;  // resident blocks per CU of the chosen wave x-sweep kernel
    int xw_occ_fr = 1;  // ... of its form that handles face rows itself (slab schedule)
    int xw_cap = 0;  // blocks per CU the wave x sweep launches (0: xw_occ)
    double* bulk_st = nullptr;   // bulk sources: T*S per value (device order), FP64
    double* bulk_ups = nullptr;  // bulk sources: U + S per value
    hipStream_t stream = nullptr;
    void* u = nullptr;

    std::vector<double> D, decay;
    bool coef_set = false;
    std::vector<uint8_t> active, reg;
    std::vector<double> fvals;  // [6][ns]
    // per-voxel Dirichlet face planes (empty: face table), FP64 in the device
    // plane order [ns][shell] (see SweepArgs::plane), and their device copy in T
    std::vector<double> hplanes;
    long long shell = 0;
    void* planeT = nullptr;
    std::vector<double> b, s;   // FP32 basis
    bool basis_fixed = false;

    // device coefficient storage, axis a in {0,1,2}
    double* den64[3] = {nullptr, nullptr, nullptr};
    double* cp64[3] = {nullptr, nullptr, nullptr};
    void* invT[3] = {nullptr, nullptr, nullptr};
    void* cpT[3] = {nullptr, nullptr, nullptr};
    void* kT = nullptr;
    void* betaT = nullptr;
    void* dvalT = nullptr;  // [7][ns]
    double* params = nullptr;  // D, decay, b, s, 1/s   [5][ns]
    double* red_partial = nullptr;
    double* red_out = nullptr;
    int red_blocks = 0;
    double pre_dt = -1.0;
    bool pre_valid = false;

    // agents (sorted by device voxel, stable in agent index)
    int nag = 0, nseg = 0;
    int* seg_off = nullptr;
    long long* seg_vox = nullptr;
    void* agA = nullptr;  // c1 (FP64), sorted [ns][nag]
    void* agB = nullptr;  // c2 (FP64)
    void* agC = nullptr;  // e2 (FP64)
    std::vector<double> h_c1, h_c2, h_e2;  // host copies of the sorted constants
    // internalised-substrate tracking: the agent count of the last agent call
    // (caller order), the caller index of each sorted agent, the net export
    // e1 = E*dt [n_in][ns] (caller order; nullptr = none), and when tracking
    // is on the per-agent totals [n_in][ns]
    int n_in = 0;
    int* ag_orig = nullptr;
    double* ag_e1 = nullptr;
    double* trk = nullptr;
    bool agent_lines = false;  // agent step stores whole 128-byte lines (bfa_set_agent_stores)
    int* long_seg = nullptr;   // segments of more than kLongSegment agents (k_agents_long)
    int nlong = 0;

    // graph of one step
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double graph_dt = -1.0;
    bool graph_valid = false;

    int Lreq[3] = {0, 0, 0};
    int Rreq = 0;
    int TXreq[3] = {0, 0, 0};
    int L[3] = {0, 0, 0};
    int R = 0;
    int TX[3] = {0, 64, 64};
    int variant = BFA_LOD_3D_BLOCKING;  // which reference solver's Dirichlet passes
    bool keep_first_dt = false;  // reference caching: coefficients of the first dt kept for later dts
    int slab = 0;  // k-planes per x/y slab pair (0: whole-volume sweeps)
    int split = 1;  // substrate groups swept on parallel streams (1 or 2)
    int runs[3] = {1, 1, 1};  // y / z sweeps: tile sub-runs per XCD (bfa_set_col_runs)

    hipStream_t aux[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr;
    void* scratch = nullptr;  // device address of the static scratch array on this device
    hipEvent_t ev_join[2] = {nullptr, nullptr};

    std::string err;
};

namespace {

int fail(bfa_env* e, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e)
        e->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIP_TRY(e, call)                                                                     \
    do {                                                                                     \
        hipError_t _st = (call);                                                             \
        if (_st != hipSuccess)                                                               \
            return fail((e), BFA_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(_st)); \
    } while (0)

#define LAUNCH_CHECK(e, what)                                                                 \
    do {                                                                                      \
        hipError_t _st = hipGetLastError();                                                   \
        if (_st != hipSuccess)                                                                \
            return fail((e), BFA_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(_st)); \
    } while (0)

template <typename P>
int dalloc(bfa_env* e, P** p, size_t bytes) {
    void* q = nullptr;
    hipError_t st = hipMalloc(&q, bytes ? bytes : 16);
    if (st != hipSuccess) {
        (void)hipGetLastError();
        return fail(e, BFA_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(st));
    }
    *p = reinterpret_cast<P*>(q);
    return BFA_OK;
}

template <typename P>
void dfree(P*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int axis_len(const bfa_env* e, int a) { return a == 0 ? e->nx : (a == 1 ? e->ny : e->nz); }

// Smallest chunk length whose chunk count fits pmax.
int pick_L(int n, int pmax) {
    for (int L : kLs)
        if ((n + L - 1) / L <= pmax) return L;
    return 0;
}

size_t lds_col(const bfa_env* e, int n, int L, int TX) {
    const int P = (n + L - 1) / L;
    // the FP32 form's copy of the values, when the kernel keeps it in LDS
    const size_t stash = e->prec == BFA_FP32 && !bfa::col_stash_in_regs(L) ? (size_t)TX * P * L : 0;
    return ((size_t)(2 * P * L + TX * P + P) + stash) * e->tsize;
}

size_t lds_x(const bfa_env* e, int R, int L) {
    const int P = (e->nx + L - 1) / L;
    const int S = e->nx | 1;
    return ((size_t)R * S + 2 * (size_t)P * L + (size_t)P * R + P) * e->tsize;
}

// Resident blocks per CU of a kernel at the given block size (at least 1).
static int occupancy_blocks(const void* fn, int block) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, 0) != hipSuccess || nb < 1) {
        (void)hipGetLastError();
        nb = 1;
    }
    return nb;
}

int choose_tiles(bfa_env* e) {
    // y and z sweeps: TX columns per block (32 or 64), P = ceil(n/L) chunk
    // rows with TX*P <= 1024 threads.
    for (int a = 1; a < 3; ++a) {
        const int n = axis_len(e, a);
        int TX = e->TXreq[a], L = e->Lreq[a];
        if (TX == 0 && L == 0) {
            // keep the register array at 16 nodes (no spills, full occupancy)
            // and narrow the block when the line needs more chunks
            const int lmax = 16;
            TX = 64;
            L = pick_L(n, 16);
            if (L == 0 || L > lmax) {
                TX = 32;
                L = pick_L(n, 32);
            }
            if (L == 0 || L > lmax) {
                TX = 16;
                L = pick_L(n, 64);
            }
            if (e->prec == BFA_FP64 && L > 8 && (n + 7) / 8 <= 64) {
                TX = 16;
                L = 8;
            }
        } else if (TX == 0) {
            const int P = (n + L - 1) / L;
            TX = P <= 16 ? 64 : (P <= 32 ? 32 : 16);
        } else if (L == 0) {
            L = pick_L(n, 1024 / TX);
        }
        if (TX != 16 && TX != 32 && TX != 64)
            return fail(e, BFA_ERR_ARG, "columns per block must be 16, 32 or 64, got %d", TX);
        if (L == 0 || std::find(std::begin(kColLs), std::end(kColLs), L) == std::end(kColLs))
            return fail(e, BFA_ERR_ARG, "axis %d of length %d: unsupported chunk length %d", a, n, L);
        if (((n + L - 1) / L) * TX > 1024)
            return fail(e, BFA_ERR_ARG, "axis %d of length %d: %d chunks of %d x %d columns exceed 1024 threads",
                        a, n, (n + L - 1) / L, L, TX);
        if (lds_col(e, n, L, TX) > (size_t)e->max_lds)
            return fail(e, BFA_ERR_ARG, "axis %d: LDS need %zu exceeds %d", a, lds_col(e, n, L, TX),
                        e->max_lds);
        e->L[a] = L;
        e->TX[a] = TX;
    }
    // x sweep.  Rreq == 0: wave-local kernel (rows of P = ceil(nx/L) <= 64
    // chunks inside one wavefront).  Rreq >= 1: LDS-tiled kernel with Rreq rows.
    const int nx = e->nx;
    if (e->Rreq == 0) {
        int L = e->Lreq[0];
        if (L == 0) {
            L = e->prec == BFA_FP32 ? 16 : 8;
            while (L < 64 && (nx + L - 1) / L > 64) L <<= 1;
        }
        if (std::find(std::begin(kXWaveLs), std::end(kXWaveLs), L) == std::end(kXWaveLs))
            return fail(e, BFA_ERR_ARG, "wave x sweep: unsupported chunk length %d", L);
        if ((nx + L - 1) / L > 64)
            return fail(e, BFA_ERR_ARG, "wave x sweep: nx=%d needs more than 64 chunks of %d", nx, L);
        e->L[0] = L;
        e->R = 0;
        return BFA_OK;
    }
    int R = e->Rreq;
    if (R < 1 || R > 64 || (R & (R - 1)))
        return fail(e, BFA_ERR_ARG, "x rows per tile must be a power of two in [1,64], got %d", R);
    int L = e->Lreq[0] ? e->Lreq[0] : pick_L(nx, std::max(1, 1024 / R));
    if (L == 0 || std::find(std::begin(kLs), std::end(kLs), L) == std::end(kLs))
        return fail(e, BFA_ERR_ARG, "x axis of length %d: no chunk length fits %d rows", nx, R);
    const int P = (nx + L - 1) / L;
    const int thr = ((P * R + 63) / 64) * 64;
    if (thr > 1024)
        return fail(e, BFA_ERR_ARG, "x tile %d rows x %d chunks exceeds 1024 threads", R, P);
    if (lds_x(e, R, L) > (size_t)e->max_lds)
        return fail(e, BFA_ERR_ARG, "x tile LDS %zu exceeds %d", lds_x(e, R, L), e->max_lds);
    e->L[0] = L;
    e->R = R;
    return BFA_OK;
}

template <typename T>
int raise_lds_limits(bfa_env* e);

int apply_tiles(bfa_env* e) {
    int rc = choose_tiles(e);
    if (rc) return rc;
    return e->prec == BFA_FP32 ? raise_lds_limits<float>(e) : raise_lds_limits<double>(e);
}

// A launch window: substrates [d0, d0+nd), lines [first, first+count).
struct Win {
    int d0, nd;
    long long first, count;
};

long long axis_lines(const bfa_env* e, int axis) {
    return axis == 0 ? (long long)e->ny * e->nz
                     : (axis == 1 ? (long long)e->nz * e->nx : (long long)e->ny * e->nx);
}

Win full_win(const bfa_env* e, int axis) { return Win{0, e->ns, 0, axis_lines(e, axis)}; }

template <typename T>
bfa::SweepArgs<T> sweep_args(const bfa_env* e, int axis) {
    bfa::SweepArgs<T> a;
    a.u = reinterpret_cast<T*>(e->u);
    a.inv = reinterpret_cast<const T*>(e->invT[axis]);
    a.cp = reinterpret_cast<const T*>(e->cpT[axis]);
    a.k = reinterpret_cast<const T*>(e->kT);
    a.beta = reinterpret_cast<const T*>(e->betaT);
    a.dval = reinterpret_cast<const T*>(e->dvalT);
    a.plane = e->hplanes.empty() ? nullptr : reinterpret_cast<const T*>(e->planeT);
    a.shell = e->shell;
    unsigned long long act = 0;
    for (int d = 0; d < e->ns; ++d)
        if (e->active[d]) act |= 1ull << d;
    // the z sweep's only Dirichlet pass is the one after it, which
    // LOD_3D_AVX256D does not make; with per-voxel planes the sweeps apply no
    // Dirichlet values and standalone face passes run between them
    // (enqueue_diffusion)
    a.active = ((axis == 2 && e->variant == BFA_LOD_3D_AVX256D) || a.plane) ? 0ull : act;
    unsigned reg = 0;
    for (int f = 0; f < 6; ++f)
        if (e->reg[f]) reg |= 1u << f;
    a.reg = reg;
    a.N = e->N;
    a.nx = e->nx;
    a.ny = e->ny;
    a.nz = e->nz;
    a.ns = e->ns;
    a.d0 = 0;
    a.first = 0;
    a.count = axis == 0 ? (long long)e->ny * e->nz
                        : (axis == 1 ? (long long)e->nz * e->nx : (long long)e->ny * e->nx);
    a.runs = e->runs[axis];
    return a;
}

template <typename T, int L>
int launch_x_L(bfa_env* e, hipStream_t st) {
    constexpr bool DEV = std::is_same<T, float>::value;
    const int R = e->R;
    int logR = 0;
    while ((1 << logR) < R) ++logR;
    const int P = (e->nx + L - 1) / L;
    const int thr = ((P * R + 63) / 64) * 64;
    const long long rows = (long long)e->ny * e->nz;
    const size_t lds = lds_x(e, R, L);
    dim3 grid((unsigned)((rows + R - 1) / R), (unsigned)e->ns);
    auto fn = bfa::k_sweep_x<T, L, DEV>;
    hipLaunchKernelGGL(fn, grid, dim3(thr), lds, st, sweep_args<T>(e, 0), R, logR);
    LAUNCH_CHECK(e, "x sweep");
    return BFA_OK;
}

template <typename T, int L>
int launch_xw_L(bfa_env* e, hipStream_t st, const Win& w) {
    constexpr bool DEV = std::is_same<T, float>::value;
    const int nx = e->nx;
    const int P = (nx + L - 1) / L;
    const int RW = 64 / P;
    const long long groups = (w.count + RW - 1) / RW;
    // one resident wave of blocks fills the device; each wavefront then
    // strides over its row groups with the next one's loads in flight
    // The slab schedule (x and y sweeps of a slab back to back) keeps the
    // rows on y and z faces in the kernel; otherwise a face pass over the
    // window's substrates follows the sweep and completes its Dirichlet pass.
    const bool fr = e->slab > 0;
    const int resident = fr ? e->xw_occ_fr : e->xw_occ;
    const int occ = e->xw_cap > 0 ? std::min(e->xw_cap, resident) : resident;
    const long long want = std::max<long long>(1, (long long)e->cus * occ / w.nd);
    const unsigned bx = (unsigned)std::max<long long>(1, std::min<long long>((groups + 3) / 4, want));
    bfa::SweepArgs<T> a = sweep_args<T>(e, 0);
    a.d0 = w.d0;
    a.first = w.first;
    a.count = w.count;
    dim3 grid(bx, (unsigned)w.nd);
    if (fr)
        hipLaunchKernelGGL((bfa::k_sweep_x_wave<T, L, DEV, true>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((bfa::k_sweep_x_wave<T, L, DEV, false>), grid, dim3(256), 0, st, a);
    LAUNCH_CHECK(e, "x sweep (wave)");
    if (!fr && a.active) {
        unsigned long long group = 0;
        for (int d = w.d0; d < w.d0 + w.nd; ++d) group |= 1ull << d;
        a.active &= group;
        if (a.active) {
            const long long work = 2ll * e->nx * e->nz + 2ll * e->nx * e->ny;  // y and z faces
            hipLaunchKernelGGL(bfa::k_dirichlet_faces<T>, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, a,
                               1);
            LAUNCH_CHECK(e, "x sweep face pass");
        }
    }
    return BFA_OK;
}

template <typename T, int L, int AXIS>
int launch_col_L(bfa_env* e, hipStream_t st, const Win& w) {
    constexpr bool DEV = std::is_same<T, float>::value;
    const int n = axis_len(e, AXIS);
    const int P = (n + L - 1) / L;
    const int TX = e->TX[AXIS];
    const size_t lds = lds_col(e, n, L, TX);
    bfa::SweepArgs<T> a = sweep_args<T>(e, AXIS);
    a.d0 = w.d0;
    a.first = w.first;
    a.count = w.count;
    dim3 grid((unsigned)((w.count + TX - 1) / TX), (unsigned)w.nd);
    if (TX == 64)
        hipLaunchKernelGGL((bfa::k_sweep_col<T, L, AXIS, DEV, 64>), grid, dim3(64 * P), lds, st, a);
    else if (TX == 32)
        hipLaunchKernelGGL((bfa::k_sweep_col<T, L, AXIS, DEV, 32>), grid, dim3(32 * P), lds, st, a);
    else
        hipLaunchKernelGGL((bfa::k_sweep_col<T, L, AXIS, DEV, 16>), grid, dim3(16 * P), lds, st, a);
    LAUNCH_CHECK(e, AXIS == 1 ? "y sweep" : "z sweep");
    return BFA_OK;
}

// Kernels whose dynamic LDS exceeds 64 KiB need the per-function limit raised
// once, outside any stream capture.
template <typename T>
int raise_lds_limits(bfa_env* e) {
    constexpr bool DEV = std::is_same<T, float>::value;
    const void* fx = nullptr;
    const void* fy = nullptr;
    const void* fz = nullptr;
#define BFA_COLFN(Lv, A)                                                       \
    (e->TX[A] == 64   ? (const void*)bfa::k_sweep_col<T, Lv, A, DEV, 64>       \
     : e->TX[A] == 32 ? (const void*)bfa::k_sweep_col<T, Lv, A, DEV, 32>       \
                      : (const void*)bfa::k_sweep_col<T, Lv, A, DEV, 16>)
#define BFA_PICK(Lv)                                                          \
    if (e->L[0] == Lv && e->R) fx = (const void*)bfa::k_sweep_x<T, Lv, DEV>;  \
    if (e->L[1] == Lv) fy = BFA_COLFN(Lv, 1);                                 \
    if (e->L[2] == Lv) fz = BFA_COLFN(Lv, 2);
    BFA_PICK(4) BFA_PICK(8) BFA_PICK(16) BFA_PICK(32) BFA_PICK(64)
    // column-only chunk length: 25 nodes make 16 chunks of a 400-node line,
    // so 32 columns fill eight whole wavefronts with no padded nodes
    if (e->L[1] == 25) fy = BFA_COLFN(25, 1);
    if (e->L[2] == 25) fz = BFA_COLFN(25, 2);
#undef BFA_PICK
#undef BFA_COLFN
    const size_t need[3] = {e->R ? lds_x(e, e->R, e->L[0]) : 0, lds_col(e, e->ny, e->L[1], e->TX[1]),
                            lds_col(e, e->nz, e->L[2], e->TX[2])};
    const void* fn[3] = {fx, fy, fz};
    for (int a = 0; a < 3; ++a)
        if (need[a] > 65536 && fn[a])
            HIP_TRY(e, hipFuncSetAttribute(fn[a], hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)need[a]));
    if (e->R == 0) {
        const void* fw = nullptr;
        const void* fwr = nullptr;
#define BFA_XW(Lv)                                                            \
    if (e->L[0] == Lv) {                                                      \
        fw = (const void*)bfa::k_sweep_x_wave<T, Lv, DEV, false>;             \
        fwr = (const void*)bfa::k_sweep_x_wave<T, Lv, DEV, true>;             \
    }
        BFA_XW(4) BFA_XW(8) BFA_XW(16) BFA_XW(20) BFA_XW(28) BFA_XW(32)
#undef BFA_XW
        e->xw_occ = fw ? occupancy_blocks(fw, 256) : 1;
        e->xw_occ_fr = fwr ? occupancy_blocks(fwr, 256) : 1;
    }
    return BFA_OK;
}

template <typename T>
int launch_sweep(bfa_env* e, int axis, hipStream_t st, const Win& w) {
    const int L = e->L[axis];
    if (axis == 0 && e->R == 0) {
        switch (L) {
            case 4: return launch_xw_L<T, 4>(e, st, w);
            case 8: return launch_xw_L<T, 8>(e, st, w);
            case 16: return launch_xw_L<T, 16>(e, st, w);
            case 20: return launch_xw_L<T, 20>(e, st, w);
            case 28: return launch_xw_L<T, 28>(e, st, w);
            case 32: return launch_xw_L<T, 32>(e, st, w);
        }
    } else if (axis == 0) {
        switch (L) {
            case 4: return launch_x_L<T, 4>(e, st);
            case 8: return launch_x_L<T, 8>(e, st);
            case 16: return launch_x_L<T, 16>(e, st);
            case 32: return launch_x_L<T, 32>(e, st);
            case 64: return launch_x_L<T, 64>(e, st);
        }
    } else if (axis == 1) {
        switch (L) {
            case 4: return launch_col_L<T, 4, 1>(e, st, w);
            case 8: return launch_col_L<T, 8, 1>(e, st, w);
            case 16: return launch_col_L<T, 16, 1>(e, st, w);
            case 25: return launch_col_L<T, 25, 1>(e, st, w);
            case 32: return launch_col_L<T, 32, 1>(e, st, w);
            case 64: return launch_col_L<T, 64, 1>(e, st, w);
        }
    } else {
        switch (L) {
            case 4: return launch_col_L<T, 4, 2>(e, st, w);
            case 8: return launch_col_L<T, 8, 2>(e, st, w);
            case 16: return launch_col_L<T, 16, 2>(e, st, w);
            case 25: return launch_col_L<T, 25, 2>(e, st, w);
            case 32: return launch_col_L<T, 32, 2>(e, st, w);
            case 64: return launch_col_L<T, 64, 2>(e, st, w);
        }
    }
    return fail(e, BFA_ERR_ARG, "no kernel for chunk length %d", L);
}

int launch_sweep_any(bfa_env* e, int axis) {
    const Win w = full_win(e, axis);
    return e->prec == BFA_FP32 ? launch_sweep<float>(e, axis, e->stream, w)
                               : launch_sweep<double>(e, axis, e->stream, w);
}

int launch_sweep_win(bfa_env* e, int axis, hipStream_t st, const Win& w) {
    return e->prec == BFA_FP32 ? launch_sweep<float>(e, axis, st, w)
                               : launch_sweep<double>(e, axis, st, w);
}

// x and y sweeps in k-slabs of e->slab planes, per substrate, alternating
// between two side streams so the x sweep of one slab overlaps the y sweep of
// another; the y sweep of a slab re-reads what its x sweep just wrote while
// it is still cache-resident.  Forks from and joins back to e->stream.
int enqueue_xy_slabs(bfa_env* e) {
    HIP_TRY(e, hipEventRecord(e->ev_fork, e->stream));
    HIP_TRY(e, hipStreamWaitEvent(e->aux[0], e->ev_fork, 0));
    HIP_TRY(e, hipStreamWaitEvent(e->aux[1], e->ev_fork, 0));
    int q = 0;
    for (int d = 0; d < e->ns; ++d)
        for (int k0 = 0; k0 < e->nz; k0 += e->slab, ++q) {
            const int k1 = std::min(e->nz, k0 + e->slab);
            hipStream_t st = e->aux[q & 1];
            const Win wx{d, 1, (long long)k0 * e->ny, (long long)(k1 - k0) * e->ny};
            const Win wy{d, 1, (long long)k0 * e->nx, (long long)(k1 - k0) * e->nx};
            int rc = launch_sweep_win(e, 0, st, wx);
            if (rc) return rc;
            rc = launch_sweep_win(e, 1, st, wy);
            if (rc) return rc;
        }
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(e, hipEventRecord(e->ev_join[i], e->aux[i]));
        HIP_TRY(e, hipStreamWaitEvent(e->stream, e->ev_join[i], 0));
    }
    return BFA_OK;
}

// The three sweeps of two substrate groups on the two side streams, so that
// each kernel's tail overlaps the other group's work.  Forks from and joins
// back to e->stream.
int launch_agents_win(bfa_env* e, hipStream_t st, int d0, int nd);

// The two substrate halves on the two side streams; with agents (a step)
// each half's agent step follows its z sweep on its stream, so one half's
// agents overlap the other half's sweeps.
int enqueue_split(bfa_env* e, bool agents) {
    HIP_TRY(e, hipEventRecord(e->ev_fork, e->stream));
    const int h = (e->ns + 1) / 2;
    for (int g = 0; g < 2; ++g) {
        hipStream_t st = e->aux[g];
        HIP_TRY(e, hipStreamWaitEvent(st, e->ev_fork, 0));
        const int d0 = g == 0 ? 0 : h, nd = g == 0 ? h : e->ns - h;
        for (int a = 0; a < 3; ++a) {
            int rc = launch_sweep_win(e, a, st, Win{d0, nd, 0, axis_lines(e, a)});
            if (rc) return rc;
        }
        if (agents) {
            int rc = launch_agents_win(e, st, d0, nd);
            if (rc) return rc;
        }
        HIP_TRY(e, hipEventRecord(e->ev_join[g], st));
    }
    for (int g = 0; g < 2; ++g) HIP_TRY(e, hipStreamWaitEvent(e->stream, e->ev_join[g], 0));
    return BFA_OK;
}

// The agent step over substrates [d0, d0 + nd) on stream st.
int launch_agents_win(bfa_env* e, hipStream_t st, int d0, int nd) {
    if (e->nseg == 0 || nd <= 0) return BFA_OK;
    const long long work = (long long)e->nseg * nd;
    const unsigned blocks = (unsigned)((work + 255) / 256);
    const double W = e->dx * e->dx * e->dx;  // voxel volume
#define BFA_AGENTS(T, LN)                                                                                     \
    hipLaunchKernelGGL((bfa::k_agents<T, LN>), dim3(blocks), dim3(256), 0, st, reinterpret_cast<T*>(e->u), e->N, \
                       e->ns, d0, nd, e->nseg, e->nag, e->seg_off, e->seg_vox,                                   \
                       reinterpret_cast<const double*>(e->agA), reinterpret_cast<const double*>(e->agB),         \
                       reinterpret_cast<const double*>(e->agC), e->params, e->trk, e->ag_orig, e->ag_e1, W)
    if (e->prec == BFA_FP32) {
        if (e->agent_lines)
            BFA_AGENTS(float, true);
        else
            BFA_AGENTS(float, false);
    } else {
        if (e->agent_lines)
            BFA_AGENTS(double, true);
        else
            BFA_AGENTS(double, false);
    }
#undef BFA_AGENTS
    LAUNCH_CHECK(e, "agents");
    if (e->nlong > 0) {
        const unsigned lblocks = (unsigned)((64ll * e->nlong * nd + 255) / 256);
#define BFA_AGENTS_LONG(T)                                                                                    \
    hipLaunchKernelGGL(bfa::k_agents_long<T>, dim3(lblocks), dim3(256), 0, st, reinterpret_cast<T*>(e->u), e->N, \
                       e->ns, d0, nd, e->nlong, e->long_seg, e->nag, e->seg_off, e->seg_vox,                    \
                       reinterpret_cast<const double*>(e->agA), reinterpret_cast<const double*>(e->agB),         \
                       reinterpret_cast<const double*>(e->agC), e->params, e->trk, e->ag_orig, e->ag_e1, W)
        if (e->prec == BFA_FP32)
            BFA_AGENTS_LONG(float);
        else
            BFA_AGENTS_LONG(double);
#undef BFA_AGENTS_LONG
        LAUNCH_CHECK(e, "agents (long segments)");
    }
    return BFA_OK;
}

int launch_agents(bfa_env* e) { return launch_agents_win(e, e->stream, 0, e->ns); }

// Substrates whose Dirichlet values are applied (dirichlet_activation_vector).
unsigned long long face_active(const bfa_env* e) {
    unsigned long long act = 0;
    for (int d = 0; d < e->ns; ++d)
        if (e->active[d]) act |= 1ull << d;
    return act;
}

// Drops the captured step graph.  Launches of it may still be queued or
// running (bfa_step only enqueues them), and its kernel arguments point into
// buffers that the caller is about to replace, so the handle's stream (which
// every side stream joins back into) is drained first.
void invalidate_graph(bfa_env* e) {
    if (e->gexec && e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    e->graph_valid = false;
}

// Upload params (D, decay, b, s, 1/s) and the Dirichlet table in T.
template <typename T>
int upload_tables(bfa_env* e) {
    const int ns = e->ns;
    std::vector<double> p(5 * (size_t)ns);
    for (int d = 0; d < ns; ++d) {
        p[d] = e->D[d];
        p[ns + d] = e->decay[d];
        p[2 * ns + d] = e->b[d];
        p[3 * ns + d] = e->s[d];
        p[4 * ns + d] = 1.0 / e->s[d];
    }
    HIP_TRY(e, hipMemcpyAsync(e->params, p.data(), p.size() * sizeof(double), hipMemcpyHostToDevice,
                              e->stream));
    std::vector<T> dv(7 * (size_t)ns);
    for (int f = 0; f < 7; ++f)
        for (int d = 0; d < ns; ++d) {
            const double val = f < 6 ? e->fvals[(size_t)f * ns + d] : 1.0;
            dv[(size_t)f * ns + d] = (T)((val - e->b[d]) / e->s[d]);
        }
    HIP_TRY(e, hipMemcpyAsync(e->dvalT, dv.data(), dv.size() * sizeof(T), hipMemcpyHostToDevice,
                              e->stream));
    if (!e->hplanes.empty()) {
        std::vector<T> pv(e->hplanes.size());
        for (int d = 0; d < ns; ++d)
            for (long long q = 0; q < e->shell; ++q) {
                const size_t o = (size_t)d * e->shell + q;
                pv[o] = (T)((e->hplanes[o] - e->b[d]) / e->s[d]);
            }
        HIP_TRY(e, hipMemcpyAsync(e->planeT, pv.data(), pv.size() * sizeof(T), hipMemcpyHostToDevice,
                                  e->stream));
        HIP_TRY(e, hipStreamSynchronize(e->stream));  // pv is released on return
    }
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return BFA_OK;
}

// FP32 basis value b of an active substrate: the Dirichlet value it is held
// at (the mean plane value with per-voxel planes; else the value of the last
// registered face, or the default 1.0), so face voxels store v near 0.
double dirichlet_basis(const bfa_env* e, int d) {
    if (!e->hplanes.empty()) {
        double acc = 0.0;
        const double* pl = e->hplanes.data() + (size_t)d * e->shell;
        for (long long q = 0; q < e->shell; ++q) acc += pl[q];
        return acc / (double)e->shell;
    }
    for (int f = 5; f >= 0; --f)
        if (e->reg[f]) return e->fvals[(size_t)f * e->ns + d];
    return 1.0;
}

template <typename T>
int run_precompute(bfa_env* e, double dt) {
    int rc = upload_tables<T>(e);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a) {
        hipLaunchKernelGGL(bfa::k_thomas<T>, dim3((e->ns + 63) / 64), dim3(64), 0, e->stream,
                           axis_len(e, a), e->ns, e->params, dt, e->dx, e->den64[a], e->cp64[a],
                           reinterpret_cast<T*>(e->invT[a]), reinterpret_cast<T*>(e->cpT[a]),
                           reinterpret_cast<T*>(e->kT), reinterpret_cast<T*>(e->betaT));
        LAUNCH_CHECK(e, "thomas precompute");
    }
    return BFA_OK;
}

int ensure_precompute(bfa_env* e, double dt) {
    if (!e->coef_set) return fail(e, BFA_ERR_STATE, "bfa_set_coefficients has not been called");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(e, BFA_ERR_ARG, "dt must be positive, got %g", dt);
    if (e->pre_valid && (e->pre_dt == dt || e->keep_first_dt)) return BFA_OK;
    invalidate_graph(e);
    int rc = e->prec == BFA_FP32 ? run_precompute<float>(e, dt) : run_precompute<double>(e, dt);
    if (rc) return rc;
    e->pre_dt = dt;
    e->pre_valid = true;
    return BFA_OK;
}

// The standalone face pass (apply_dirichlet_boundaries_conditions) on the
// handle's stream, with the table or the per-voxel planes.
int launch_faces(bfa_env* e) {
    const long long work = 2ll * e->ny * e->nz + 2ll * e->nx * e->nz + 2ll * e->nx * e->ny;
    const unsigned blocks = (unsigned)((work + 255) / 256);
    if (e->prec == BFA_FP32) {
        bfa::SweepArgs<float> a = sweep_args<float>(e, 0);
        a.active = face_active(e);
        hipLaunchKernelGGL(bfa::k_dirichlet_faces<float>, dim3(blocks), dim3(256), 0, e->stream, a, 0);
    } else {
        bfa::SweepArgs<double> a = sweep_args<double>(e, 0);
        a.active = face_active(e);
        hipLaunchKernelGGL(bfa::k_dirichlet_faces<double>, dim3(blocks), dim3(256), 0, e->stream, a, 0);
    }
    LAUNCH_CHECK(e, "dirichlet faces");
    return BFA_OK;
}

bool split_schedule(const bfa_env* e) { return e->split > 1 && e->ns > 1 && e->slab == 0; }

int enqueue_diffusion(bfa_env* e) {
    if (!e->hplanes.empty()) {
        // per-voxel Dirichlet values: the reference's separate face passes
        // around plain sweeps (BioFVM/BioFVM_solvers.cpp:345,682,731,778)
        const bool any = face_active(e) != 0;
        for (int a = 0; a < 3; ++a) {
            int rc = any ? launch_faces(e) : BFA_OK;
            if (rc == BFA_OK) rc = launch_sweep_any(e, a);
            if (rc) return rc;
        }
        return (any && e->variant == BFA_LOD_3D_BLOCKING) ? launch_faces(e) : BFA_OK;
    }
    if (split_schedule(e)) return enqueue_split(e, false);
    if (e->slab > 0 && e->R == 0) {
        int rc = enqueue_xy_slabs(e);
        if (rc) return rc;
        return launch_sweep_any(e, 2);
    }
    for (int a = 0; a < 3; ++a) {
        int rc = launch_sweep_any(e, a);
        if (rc) return rc;
    }
    return BFA_OK;
}

int enqueue_step(bfa_env* e) {
    // the substrate-halves schedule runs each half's agents on its stream
    if (e->hplanes.empty() && split_schedule(e)) return enqueue_split(e, true);
    int rc = enqueue_diffusion(e);
    if (rc) return rc;
    return launch_agents(e);
}

int build_graph(bfa_env* e, double dt) {
    invalidate_graph(e);
    HIP_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(e);
    hipGraph_t g = nullptr;
    hipError_t st = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (st != hipSuccess) return fail(e, BFA_ERR_HIP, "graph capture failed: %s", hipGetErrorString(st));
    e->graph = g;
    st = hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0);
    if (st != hipSuccess) {
        invalidate_graph(e);
        return fail(e, BFA_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(st));
    }
    e->graph_dt = dt;
    e->graph_valid = true;
    return BFA_OK;
}

int reduce(bfa_env* e, const void* x, bool f32, long long count, long long stride, long long dstride,
           int mode, const double* ref_dev, double* host_out) {
    const int nb = e->red_blocks;
    dim3 grid(nb, e->ns);
    if (f32) {
        if (mode == 0)
            hipLaunchKernelGGL((bfa::k_reduce_partial<float, 0>), grid, dim3(256), 0, e->stream,
                               (const float*)x, count, stride, dstride, ref_dev, e->red_partial);
        else
            hipLaunchKernelGGL((bfa::k_reduce_partial<float, 1>), grid, dim3(256), 0, e->stream,
                               (const float*)x, count, stride, dstride, ref_dev, e->red_partial);
    } else {
        if (mode == 0)
            hipLaunchKernelGGL((bfa::k_reduce_partial<double, 0>), grid, dim3(256), 0, e->stream,
                               (const double*)x, count, stride, dstride, ref_dev, e->red_partial);
        else
            hipLaunchKernelGGL((bfa::k_reduce_partial<double, 1>), grid, dim3(256), 0, e->stream,
                               (const double*)x, count, stride, dstride, ref_dev, e->red_partial);
    }
    LAUNCH_CHECK(e, "reduce partial");
    if (mode == 0)
        hipLaunchKernelGGL(bfa::k_reduce_final<0>, dim3(e->ns), dim3(64), 0, e->stream,
                           e->red_partial, nb, e->red_out);
    else
        hipLaunchKernelGGL(bfa::k_reduce_final<1>, dim3(e->ns), dim3(64), 0, e->stream,
                           e->red_partial, nb, e->red_out);
    LAUNCH_CHECK(e, "reduce final");
    HIP_TRY(e, hipMemcpyAsync(host_out, e->red_out, e->ns * sizeof(double), hipMemcpyDeviceToHost,
                              e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return BFA_OK;
}

int scratch_setup(bfa_env* e);  // biofvm_noalloc.inc

// A new agent list of another length has other agent identities: the
// tracked totals (indexed by agent) no longer apply and tracking stops.
void drop_tracking_if_resized(bfa_env* e, int n) {
    if (e->trk && n != e->n_in) dfree(e->trk);
}

}  // namespace

extern "C" {

const char* bfa_last_error(const bfa_env* env) {
    return env ? env->err.c_str() : g_create_error.c_str();
}

int bfa_create(bfa_env** out, int nx, int ny, int nz, int ns, double dx, int precision) {
    if (!out) return fail(nullptr, BFA_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1 || nx > kMaxEdge || ny > kMaxEdge || nz > kMaxEdge)
        return fail(nullptr, BFA_ERR_ARG, "grid %dx%dx%d outside [1,%d] per axis", nx, ny, nz, kMaxEdge);
    if (ns < 1 || ns > kMaxSubstrates)
        return fail(nullptr, BFA_ERR_ARG, "ns=%d outside [1,%d]", ns, kMaxSubstrates);
    if (!(dx > 0.0) || !std::isfinite(dx)) return fail(nullptr, BFA_ERR_ARG, "dx must be positive");
    if (precision != BFA_FP64 && precision != BFA_FP32)
        return fail(nullptr, BFA_ERR_ARG, "unknown precision %d", precision);
    bfa_env* e = new bfa_env();
    e->nx = nx;
    e->ny = ny;
    e->nz = nz;
    e->ns = ns;
    e->dx = dx;
    e->prec = precision;
    e->N = (long long)nx * ny * nz;
    e->tsize = precision == BFA_FP32 ? 4 : 8;
    e->active.assign(ns, 0);
    e->reg.assign(6, 0);
    e->fvals.assign(6 * (size_t)ns, 1.0);
    e->b.assign(ns, 0.0);
    e->s.assign(ns, 1.0);
    auto bail = [&](int rc) {
        g_create_error = e->err;
        bfa_destroy(e);
        return rc;
    };
    hipError_t st = hipGetDevice(&e->device);
    if (st != hipSuccess) {
        fail(e, BFA_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(st));
        return bail(BFA_ERR_HIP);
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) == hipSuccess &&
        lds > 0)
        e->max_lds = lds;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device) == hipSuccess &&
        cus > 0)
        e->cus = cus;
    st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (st != hipSuccess) {
        fail(e, BFA_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(st));
        return bail(BFA_ERR_HIP);
    }
    for (int i = 0; i < 2; ++i) {
        st = hipStreamCreateWithFlags(&e->aux[i], hipStreamNonBlocking);
        if (st == hipSuccess) st = hipEventCreateWithFlags(&e->ev_join[i], hipEventDisableTiming);
        if (st != hipSuccess) {
            fail(e, BFA_ERR_HIP, "stream/event creation failed: %s", hipGetErrorString(st));
            return bail(BFA_ERR_HIP);
        }
    }
    st = hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming);
    if (st != hipSuccess) {
        fail(e, BFA_ERR_HIP, "event creation failed: %s", hipGetErrorString(st));
        return bail(BFA_ERR_HIP);
    }
    // the field plus a zeroed slack that row-chunk loads may read past the end
    int rc = dalloc(e, &e->u, ((size_t)e->N * ns + kFieldSlack) * e->tsize);
    if (rc) return bail(rc);
    st = hipMemsetAsync(e->u, 0, ((size_t)e->N * ns + kFieldSlack) * e->tsize, e->stream);
    if (st != hipSuccess) {
        fail(e, BFA_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(st));
        return bail(BFA_ERR_HIP);
    }
    for (int a = 0; a < 3; ++a) {
        const size_t n = (size_t)axis_len(e, a) * ns;
        if ((rc = dalloc(e, &e->den64[a], n * 8)) || (rc = dalloc(e, &e->cp64[a], n * 8)) ||
            (rc = dalloc(e, &e->invT[a], n * e->tsize)) || (rc = dalloc(e, &e->cpT[a], n * e->tsize)))
            return bail(rc);
    }
    if ((rc = dalloc(e, &e->kT, 2 * (size_t)ns * e->tsize)) || (rc = dalloc(e, &e->betaT, ns * e->tsize)) ||
        (rc = dalloc(e, &e->dvalT, 7 * (size_t)ns * e->tsize)) ||
        (rc = dalloc(e, &e->params, 5 * (size_t)ns * 8)))
        return bail(rc);
    {
        // neutral parameters until the first upload / precompute: b = 0, s = 1
        std::vector<double> p0(5 * (size_t)ns, 0.0);
        for (int d = 0; d < ns; ++d) p0[3 * ns + d] = p0[4 * ns + d] = 1.0;
        st = hipMemcpy(e->params, p0.data(), p0.size() * 8, hipMemcpyHostToDevice);
        if (st != hipSuccess) {
            fail(e, BFA_ERR_HIP, "params init failed: %s", hipGetErrorString(st));
            return bail(BFA_ERR_HIP);
        }
    }
    e->red_blocks = (int)std::min<long long>(512, std::max<long long>(1, e->N / 8192));
    if ((rc = dalloc(e, &e->red_partial, (size_t)e->red_blocks * ns * 8)) ||
        (rc = dalloc(e, &e->red_out, (size_t)ns * 8)))
        return bail(rc);
    if ((rc = apply_tiles(e))) return bail(rc);
    if ((rc = scratch_setup(e))) return bail(rc);
    st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) {
        fail(e, BFA_ERR_HIP, "initial sync failed: %s", hipGetErrorString(st));
        return bail(BFA_ERR_HIP);
    }
    *out = e;
    return BFA_OK;
}

void bfa_destroy(bfa_env* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    invalidate_graph(e);
    dfree(e->u);
    for (int a = 0; a < 3; ++a) {
        dfree(e->den64[a]);
        dfree(e->cp64[a]);
        dfree(e->invT[a]);
        dfree(e->cpT[a]);
    }
    dfree(e->kT);
    dfree(e->betaT);
    dfree(e->dvalT);
    dfree(e->planeT);
    dfree(e->params);
    dfree(e->red_partial);
    dfree(e->red_out);
    dfree(e->seg_off);
    dfree(e->seg_vox);
    dfree(e->agA);
    dfree(e->agB);
    dfree(e->agC);
    dfree(e->ag_orig);
    dfree(e->ag_e1);
    dfree(e->trk);
    dfree(e->long_seg);
    dfree(e->bulk_st);
    dfree(e->bulk_ups);
    for (int i = 0; i < 2; ++i) {
        if (e->aux[i]) (void)hipStreamSynchronize(e->aux[i]);
        if (e->aux[i]) (void)hipStreamDestroy(e->aux[i]);
        if (e->ev_join[i]) (void)hipEventDestroy(e->ev_join[i]);
    }
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int bfa_set_coefficients(bfa_env* e, const double* D, const double* decay) {
    if (!e || !D || !decay) return fail(e, BFA_ERR_ARG, "NULL argument");
    for (int d = 0; d < e->ns; ++d)
        if (!std::isfinite(D[d]) || !std::isfinite(decay[d]) || D[d] < 0 || decay[d] < 0)
            return fail(e, BFA_ERR_ARG, "substrate %d: D=%g decay=%g must be finite and >= 0", d, D[d],
                        decay[d]);
    e->D.assign(D, D + e->ns);
    e->decay.assign(decay, decay + e->ns);
    e->coef_set = true;
    e->pre_valid = false;
    invalidate_graph(e);
    return BFA_OK;
}

int bfa_set_dirichlet(bfa_env* e, const uint8_t* active, const uint8_t* face_registered,
                      const double* face_values) {
    if (!e || !active || !face_registered || !face_values) return fail(e, BFA_ERR_ARG, "NULL argument");
    e->active.assign(active, active + e->ns);
    for (auto& a : e->active) a = a ? 1 : 0;
    e->reg.assign(face_registered, face_registered + 6);
    for (auto& r : e->reg) r = r ? 1 : 0;
    e->fvals.assign(face_values, face_values + 6 * (size_t)e->ns);
    e->pre_valid = false;
    invalidate_graph(e);
    return BFA_OK;
}

int bfa_set_dirichlet_planes(bfa_env* e, const double* planes) {
    if (!e) return BFA_ERR_ARG;
    const int nx = e->nx, ny = e->ny, nz = e->nz, ns = e->ns;
    if (!planes) {
        invalidate_graph(e);
        e->hplanes.clear();
        e->hplanes.shrink_to_fit();
        dfree(e->planeT);
        e->pre_valid = false;
        return BFA_OK;
    }
    const long long nyz = (long long)ny * nz, nxz = (long long)nx * nz, nxy = (long long)nx * ny;
    const long long shell = 2 * (nyz + nxz + nxy);
    // input: reference voxel order per plane, substrate fastest:
    //   xmin/xmax [j][k][d], ymin/ymax [i][k][d], zmin/zmax [i][j][d]
    // device: [d][shell] with xmin/xmax [k][j], ymin/ymax [k][i], zmin/zmax [j][i]
    const long long in_off[6] = {0, nyz, 2 * nyz, 2 * nyz + nxz, 2 * (nyz + nxz), 2 * (nyz + nxz) + nxy};
    std::vector<double> h((size_t)shell * ns);
    for (int f = 0; f < 6; ++f) {
        const int A = f < 2 ? ny : nx;        // slow index of the input plane
        const int B = f < 4 ? nz : ny;        // fast index
        for (int a = 0; a < A; ++a)
            for (int bq = 0; bq < B; ++bq) {
                const long long qin = in_off[f] + (long long)a * B + bq;
                // device index: x faces k*ny + j (a = j, bq = k); y faces
                // k*nx + i (a = i, bq = k); z faces j*nx + i (a = i, bq = j)
                const long long qdev = in_off[f] + (f < 4 ? (long long)bq * A + a : (long long)bq * nx + a);
                for (int d = 0; d < ns; ++d) {
                    const double v = planes[(size_t)qin * ns + d];
                    if (!std::isfinite(v))
                        return fail(e, BFA_ERR_ARG, "Dirichlet plane %d entry %lld substrate %d is not finite", f,
                                    qin, d);
                    h[(size_t)d * shell + qdev] = v;
                }
            }
    }
    // a voxel on several faces has one value (dirichlet_value_vectors[voxel])
    auto at = [&](int i, int j, int k, int d, int f) {
        long long q = 0;
        if (f < 2) q = in_off[f] + (long long)k * ny + j;
        else if (f < 4) q = in_off[f] + (long long)k * nx + i;
        else q = in_off[f] + (long long)j * nx + i;
        return h[(size_t)d * shell + q];
    };
    for (int d = 0; d < ns; ++d)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; i += (j == 0 || j == ny - 1 || k == 0 || k == nz - 1) ? 1 : std::max(1, nx - 1)) {
                    int faces[6], nf = 0;
                    if (i == 0) faces[nf++] = 0;
                    if (i == nx - 1) faces[nf++] = 1;
                    if (j == 0) faces[nf++] = 2;
                    if (j == ny - 1) faces[nf++] = 3;
                    if (k == 0) faces[nf++] = 4;
                    if (k == nz - 1) faces[nf++] = 5;
                    for (int q = 1; q < nf; ++q)
                        if (at(i, j, k, d, faces[q]) != at(i, j, k, d, faces[0]))
                            return fail(e, BFA_ERR_ARG,
                                        "voxel (%d,%d,%d) substrate %d has different values in Dirichlet planes %d "
                                        "and %d (one value per voxel)", i, j, k, d, faces[0], faces[q]);
                }
    invalidate_graph(e);
    if (!e->planeT || e->shell != shell) {
        dfree(e->planeT);
        int rc = dalloc(e, &e->planeT, (size_t)shell * ns * e->tsize);
        if (rc) return rc;
    }
    e->shell = shell;
    e->hplanes.swap(h);
    e->pre_valid = false;
    return BFA_OK;
}

int bfa_set_fp32_basis(bfa_env* e, const double* b, const double* s) {
    if (!e || !b || !s) return fail(e, BFA_ERR_ARG, "NULL argument");
    for (int d = 0; d < e->ns; ++d)
        if (!std::isfinite(b[d]) || !std::isfinite(s[d]) || s[d] == 0.0)
            return fail(e, BFA_ERR_ARG, "substrate %d: basis b=%g s=%g invalid", d, b[d], s[d]);
    e->b.assign(b, b + e->ns);
    e->s.assign(s, s + e->ns);
    e->basis_fixed = true;
    return BFA_OK;
}

int bfa_get_fp32_basis(const bfa_env* e, double* b, double* s) {
    if (!e || !b || !s) return BFA_ERR_ARG;
    std::copy(e->b.begin(), e->b.end(), b);
    std::copy(e->s.begin(), e->s.end(), s);
    return BFA_OK;
}

int bfa_upload(bfa_env* e, const double* field, int layout) {
    if (!e || !field) return fail(e, BFA_ERR_ARG, "NULL argument");
    if (layout != BFA_LAYOUT_REF && layout != BFA_LAYOUT_SOA)
        return fail(e, BFA_ERR_ARG, "unknown layout %d", layout);
    const size_t total = (size_t)e->N * e->ns;
    double* stage = nullptr;
    const bool direct = (e->prec == BFA_FP64 && layout == BFA_LAYOUT_SOA);
    if (direct) {
        HIP_TRY(e, copy_h2d(e->u, field, total * 8, e->stream));
        return BFA_OK;
    }
    if (e->prec == BFA_FP32 && e->basis_fixed) {
        // the basis is known before the field is read: the host copy threads
        // form the FP32 deviations, and half the bytes cross PCIe
        e->pre_valid = false;
        std::vector<double> p(5 * (size_t)e->ns, 0.0), rs(e->ns);
        for (int d = 0; d < e->ns; ++d) {
            rs[d] = 1.0 / e->s[d];
            p[d] = e->coef_set ? e->D[d] : 0.0;
            p[e->ns + d] = e->coef_set ? e->decay[d] : 0.0;
            p[2 * e->ns + d] = e->b[d];
            p[3 * e->ns + d] = e->s[d];
            p[4 * e->ns + d] = rs[d];
        }
        hipError_t st = hipMemcpyAsync(e->params, p.data(), p.size() * 8, hipMemcpyHostToDevice, e->stream);
        if (st != hipSuccess) return fail(e, BFA_ERR_HIP, "params copy failed: %s", hipGetErrorString(st));
        const bool ref = layout == BFA_LAYOUT_REF;
        float* st32 = nullptr;
        if (ref) {
            int rc = dalloc(e, &st32, total * 4);
            if (rc) return rc;
        }
        st = copy_h2d_to_f32(ref ? st32 : (float*)e->u, field, total, e->b.data(), rs.data(), e->ns, e->N, ref,
                             e->stream);
        if (st == hipSuccess && ref) {
            dim3 grid((unsigned)(((long long)e->nz * e->ns + 31) / 32), (unsigned)((e->nx + 31) / 32),
                      (unsigned)e->ny);
            hipLaunchKernelGGL(bfa::k_ref32_to_soa, grid, dim3(32, 8), 0, e->stream, st32, (float*)e->u, e->nx,
                               e->ny, e->nz, e->ns);
            st = hipGetLastError();
        }
        const hipError_t ss = hipStreamSynchronize(e->stream);
        if (st == hipSuccess) st = ss;
        dfree(st32);
        if (st != hipSuccess) return fail(e, BFA_ERR_HIP, "upload failed: %s", hipGetErrorString(st));
        return BFA_OK;
    }
    int rc = dalloc(e, &stage, total * 8);
    if (rc) return rc;
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(e->stream);
        dfree(stage);
        return code;
    };
    hipError_t st = copy_h2d(stage, field, total * 8, e->stream);
    if (st != hipSuccess)
        return cleanup(fail(e, BFA_ERR_HIP, "upload copy failed: %s", hipGetErrorString(st)));
    // basis for the FP32 deviation form
    if (e->prec == BFA_FP32 && !e->basis_fixed) {
        const long long stride = layout == BFA_LAYOUT_REF ? e->ns : 1;
        const long long dstride = layout == BFA_LAYOUT_REF ? 1 : e->N;
        std::vector<double> sum(e->ns), mx(e->ns);
        if ((rc = reduce(e, stage, false, e->N, stride, dstride, 0, nullptr, sum.data())))
            return cleanup(rc);
        for (int d = 0; d < e->ns; ++d) e->b[d] = e->active[d] ? dirichlet_basis(e, d) : sum[d] / (double)e->N;
        std::vector<double> bdev(e->b);
        double* bref = nullptr;
        if ((rc = dalloc(e, &bref, e->ns * 8))) return cleanup(rc);
        st = hipMemcpyAsync(bref, bdev.data(), e->ns * 8, hipMemcpyHostToDevice, e->stream);
        if (st == hipSuccess) rc = reduce(e, stage, false, e->N, stride, dstride, 1, bref, mx.data());
        (void)hipStreamSynchronize(e->stream);
        dfree(bref);
        if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "basis copy failed"));
        if (rc) return cleanup(rc);
        for (int d = 0; d < e->ns; ++d) {
            int ex = 0;
            if (mx[d] > 0.0 && std::isfinite(mx[d])) {
                std::frexp(mx[d], &ex);
                e->s[d] = std::ldexp(1.0, ex);
            } else {
                e->s[d] = 1.0;
            }
        }
    }
    if (e->prec == BFA_FP32) e->pre_valid = false;  // beta and the Dirichlet table follow the basis
    // params buffer: b at [2ns], 1/s at [4ns] (and s at [3ns]) for the conversion
    std::vector<double> p(5 * (size_t)e->ns, 0.0);
    for (int d = 0; d < e->ns; ++d) {
        p[d] = e->coef_set ? e->D[d] : 0.0;
        p[e->ns + d] = e->coef_set ? e->decay[d] : 0.0;
        p[2 * e->ns + d] = e->b[d];
        p[3 * e->ns + d] = e->s[d];
        p[4 * e->ns + d] = 1.0 / e->s[d];
    }
    st = hipMemcpyAsync(e->params, p.data(), p.size() * 8, hipMemcpyHostToDevice, e->stream);
    if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "params copy failed"));
    const double* bd = e->params + 2 * e->ns;
    const double* rsd = e->params + 4 * e->ns;
    if (layout == BFA_LAYOUT_REF) {
        dim3 grid((unsigned)(((long long)e->nz * e->ns + 31) / 32), (unsigned)((e->nx + 31) / 32),
                  (unsigned)e->ny);
        if (e->prec == BFA_FP32)
            hipLaunchKernelGGL(bfa::k_ref_to_soa<float>, grid, dim3(32, 8), 0, e->stream, stage,
                               (float*)e->u, e->nx, e->ny, e->nz, e->ns, bd, rsd);
        else
            hipLaunchKernelGGL(bfa::k_ref_to_soa<double>, grid, dim3(32, 8), 0, e->stream, stage,
                               (double*)e->u, e->nx, e->ny, e->nz, e->ns, bd, rsd);
    } else {
        dim3 grid((unsigned)std::min<long long>(4096, (e->N + 255) / 256), (unsigned)e->ns);
        hipLaunchKernelGGL(bfa::k_soa_from_f64<float>, grid, dim3(256), 0, e->stream, stage,
                           (float*)e->u, e->N, bd, rsd);
    }
    st = hipGetLastError();
    if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "upload conversion launch failed: %s",
                                              hipGetErrorString(st)));
    st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "upload failed: %s", hipGetErrorString(st)));
    return cleanup(BFA_OK);
}

int bfa_fill(bfa_env* e, const double* values) {
    if (!e || !values) return fail(e, BFA_ERR_ARG, "NULL argument");
    const int ns = e->ns;
    for (int d = 0; d < ns; ++d)
        if (!std::isfinite(values[d])) return fail(e, BFA_ERR_ARG, "value of substrate %d is not finite", d);
    if (e->prec == BFA_FP32 && !e->basis_fixed) {
        for (int d = 0; d < ns; ++d) {
            const double b = e->active[d] ? dirichlet_basis(e, d) : values[d];
            e->b[d] = b;
            const double dev = std::fabs(values[d] - b);
            int ex = 0;
            if (dev > 0.0 && std::isfinite(dev)) {
                std::frexp(dev, &ex);
                e->s[d] = std::ldexp(1.0, ex);
            } else {
                e->s[d] = 1.0;
            }
        }
        e->pre_valid = false;  // beta and the Dirichlet table follow the basis
    }
    // params: D, decay, b, s, 1/s as an upload leaves them; then the values
    std::vector<double> p(6 * (size_t)ns, 0.0);
    for (int d = 0; d < ns; ++d) {
        p[d] = e->coef_set ? e->D[d] : 0.0;
        p[ns + d] = e->coef_set ? e->decay[d] : 0.0;
        p[2 * ns + d] = e->b[d];
        p[3 * ns + d] = e->s[d];
        p[4 * ns + d] = 1.0 / e->s[d];
        p[5 * ns + d] = values[d];
    }
    double* vals = nullptr;
    int rc = dalloc(e, &vals, (size_t)ns * 8);
    if (rc) return rc;
    hipError_t st = hipMemcpyAsync(e->params, p.data(), 5 * (size_t)ns * 8, hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess)
        st = hipMemcpyAsync(vals, p.data() + 5 * (size_t)ns, (size_t)ns * 8, hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess) {
        dim3 grid((unsigned)std::min<long long>(4096, (e->N + 255) / 256), (unsigned)ns);
        const double* bd = e->params + 2 * ns;
        const double* rsd = e->params + 4 * ns;
        if (e->prec == BFA_FP32)
            hipLaunchKernelGGL(bfa::k_fill_const<float>, grid, dim3(256), 0, e->stream, (float*)e->u, e->N,
                               vals, bd, rsd);
        else
            hipLaunchKernelGGL(bfa::k_fill_const<double>, grid, dim3(256), 0, e->stream, (double*)e->u,
                               e->N, vals, bd, rsd);
        st = hipGetLastError();
    }
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    (void)hipStreamSynchronize(e->stream);
    dfree(vals);
    if (st != hipSuccess) return fail(e, BFA_ERR_HIP, "fill failed: %s", hipGetErrorString(st));
    return BFA_OK;
}

int bfa_download(bfa_env* e, double* field, int layout) {
    if (!e || !field) return fail(e, BFA_ERR_ARG, "NULL argument");
    if (layout != BFA_LAYOUT_REF && layout != BFA_LAYOUT_SOA)
        return fail(e, BFA_ERR_ARG, "unknown layout %d", layout);
    const size_t total = (size_t)e->N * e->ns;
    if (e->prec == BFA_FP64 && layout == BFA_LAYOUT_SOA) {
        HIP_TRY(e, copy_d2h(field, e->u, total * 8, e->stream));
        return BFA_OK;
    }
    if (e->prec == BFA_FP32) {
        // the FP32 deviations cross PCIe (half the bytes) and the host copy
        // threads expand them to u = b + s*v
        const bool ref = layout == BFA_LAYOUT_REF;
        float* st32 = nullptr;
        hipError_t st = hipSuccess;
        if (ref) {
            int rc = dalloc(e, &st32, total * 4);
            if (rc) return rc;
            dim3 grid((unsigned)(((long long)e->nz * e->ns + 31) / 32), (unsigned)((e->nx + 31) / 32),
                      (unsigned)e->ny);
            hipLaunchKernelGGL(bfa::k_soa_to_ref32, grid, dim3(32, 8), 0, e->stream, (const float*)e->u, st32,
                               e->nx, e->ny, e->nz, e->ns);
            st = hipGetLastError();
        }
        if (st == hipSuccess)
            st = copy_d2h_from_f32(field, ref ? st32 : (const float*)e->u, total, e->b.data(), e->s.data(), e->ns,
                                   e->N, ref, e->stream);
        (void)hipStreamSynchronize(e->stream);
        dfree(st32);
        if (st != hipSuccess) return fail(e, BFA_ERR_HIP, "download failed: %s", hipGetErrorString(st));
        return BFA_OK;
    }
    double* stage = nullptr;
    int rc = dalloc(e, &stage, total * 8);
    if (rc) return rc;
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(e->stream);
        dfree(stage);
        return code;
    };
    // b and s for the back-conversion live in a small scratch copy
    std::vector<double> bs(2 * (size_t)e->ns);
    for (int d = 0; d < e->ns; ++d) {
        bs[d] = e->b[d];
        bs[e->ns + d] = e->s[d];
    }
    double* dbs = nullptr;
    if ((rc = dalloc(e, &dbs, bs.size() * 8))) return cleanup(rc);
    hipError_t st = hipMemcpyAsync(dbs, bs.data(), bs.size() * 8, hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess) {  // FP64, reference layout
        dim3 grid((unsigned)(((long long)e->nz * e->ns + 31) / 32), (unsigned)((e->nx + 31) / 32),
                  (unsigned)e->ny);
        hipLaunchKernelGGL(bfa::k_soa_to_ref<double>, grid, dim3(32, 8), 0, e->stream, (const double*)e->u,
                           stage, e->nx, e->ny, e->nz, e->ns, dbs, dbs + e->ns);
        st = hipGetLastError();
    }
    if (st == hipSuccess) st = copy_d2h(field, stage, total * 8, e->stream);
    dfree(dbs);
    if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "download failed: %s", hipGetErrorString(st)));
    return cleanup(BFA_OK);
}

// Lists the segments of more than kLongSegment agents (k_agents_long applies
// them), on the device, after every agent update.  The per-segment kernel
// skips those segments, so on failure the agents are dropped (none rather
// than some applied) and the error returned.
int build_long_segments(bfa_env* e) {
    e->nlong = 0;
    dfree(e->long_seg);
    if (e->nseg == 0) return BFA_OK;
    int* cnt = nullptr;
    int rc;
    if ((rc = dalloc(e, &e->long_seg, (size_t)e->nseg * sizeof(int))) || (rc = dalloc(e, &cnt, sizeof(int)))) {
        e->nseg = e->nag = 0;
        return rc;
    }
    int h = 0;
    hipError_t st = hipMemsetAsync(cnt, 0, sizeof(int), e->stream);
    if (st == hipSuccess) {
        hipLaunchKernelGGL(bfa::k_long_segments, dim3((unsigned)((e->nseg + 255) / 256)), dim3(256), 0, e->stream,
                           e->nseg, e->seg_off, e->long_seg, cnt);
        st = hipGetLastError();
    }
    if (st == hipSuccess) st = hipMemcpyAsync(&h, cnt, sizeof(int), hipMemcpyDeviceToHost, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    dfree(cnt);
    if (st != hipSuccess) {
        e->nseg = e->nag = 0;
        return fail(e, BFA_ERR_HIP, "long agent segments: %s", hipGetErrorString(st));
    }
    e->nlong = h;
    if (h == 0) dfree(e->long_seg);
    return BFA_OK;
}

int bfa_set_agents(bfa_env* e, int n, const int32_t* voxel, const uint8_t* active, const double* c1,
                   const double* c2, const double* e2) {
    if (!e) return BFA_ERR_ARG;
    if (n < 0) return fail(e, BFA_ERR_ARG, "negative agent count");
    if (n > 0 && (!voxel || !c1 || !c2 || !e2)) return fail(e, BFA_ERR_ARG, "NULL argument");
    invalidate_graph(e);
    dfree(e->seg_off);
    dfree(e->seg_vox);
    dfree(e->agA);
    dfree(e->agB);
    dfree(e->agC);
    dfree(e->ag_orig);
    dfree(e->long_seg);
    e->nlong = 0;
    drop_tracking_if_resized(e, n);
    // host constants carry no e1: tracking kept on for the same agents keeps
    // the e1 it had, else bfa_track_internalized sets it
    if (!e->trk) dfree(e->ag_e1);
    e->n_in = n;
    e->nag = e->nseg = 0;
    e->h_c1.clear();
    e->h_c2.clear();
    e->h_e2.clear();
    if (n == 0) return BFA_OK;
    const int ns = e->ns;
    const long long nvox = e->N;
    // keep active agents with a valid voxel, keyed by device voxel index
    std::vector<long long> key;
    std::vector<int> idx;
    key.reserve(n);
    idx.reserve(n);
    for (int a = 0; a < n; ++a) {
        if (active && !active[a]) continue;
        const long long v = voxel[a];
        if (v < 0) continue;
        if (v >= nvox) return fail(e, BFA_ERR_ARG, "agent %d: voxel %lld out of range", a, v);
        const long long kk = v % e->nz, jj = (v / e->nz) % e->ny, ii = v / ((long long)e->nz * e->ny);
        key.push_back((kk * e->ny + jj) * e->nx + ii);
        idx.push_back(a);
    }
    const int m = (int)idx.size();
    if (m == 0) return BFA_OK;
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
    std::vector<int> off;
    std::vector<long long> svox;
    for (int r = 0; r < m; ++r)
        if (r == 0 || key[order[r]] != key[order[r - 1]]) {
            off.push_back(r);
            svox.push_back(key[order[r]]);
        }
    off.push_back(m);
    e->h_c1.resize((size_t)ns * m);
    e->h_c2.resize((size_t)ns * m);
    e->h_e2.resize((size_t)ns * m);
    for (int r = 0; r < m; ++r) {
        const int a = idx[order[r]];
        for (int d = 0; d < ns; ++d) {
            const size_t src = (size_t)a * ns + d, dst = (size_t)d * m + r;
            e->h_c1[dst] = c1[src];
            e->h_c2[dst] = c2[src];
            e->h_e2[dst] = e2[src];
        }
    }
    const int nseg = (int)svox.size();
    int rc;
    if ((rc = dalloc(e, &e->seg_off, off.size() * sizeof(int))) ||
        (rc = dalloc(e, &e->seg_vox, svox.size() * sizeof(long long))))
        return rc;
    HIP_TRY(e, hipMemcpy(e->seg_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->seg_vox, svox.data(), svox.size() * sizeof(long long), hipMemcpyHostToDevice));
    const size_t cnt = (size_t)ns * m;
    if ((rc = dalloc(e, &e->agA, cnt * 8)) || (rc = dalloc(e, &e->agB, cnt * 8)) ||
        (rc = dalloc(e, &e->agC, cnt * 8)))
        return rc;
    HIP_TRY(e, hipMemcpy(e->agA, e->h_c1.data(), cnt * 8, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->agB, e->h_c2.data(), cnt * 8, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->agC, e->h_e2.data(), cnt * 8, hipMemcpyHostToDevice));
    std::vector<int> orig(m);
    for (int r = 0; r < m; ++r) orig[r] = idx[order[r]];
    if ((rc = dalloc(e, &e->ag_orig, (size_t)m * sizeof(int)))) return rc;
    HIP_TRY(e, hipMemcpy(e->ag_orig, orig.data(), (size_t)m * sizeof(int), hipMemcpyHostToDevice));
    e->nag = m;
    e->nseg = nseg;
    return build_long_segments(e);
}

int bfa_track_internalized(bfa_env* e, int enable, const double* initial, const double* e1) {
    if (!e) return BFA_ERR_ARG;
    invalidate_graph(e);
    if (!enable) {
        dfree(e->trk);
        return BFA_OK;
    }
    const int n = e->n_in;
    if (n <= 0) return fail(e, BFA_ERR_STATE, "no agents: call bfa_set_agents or bfa_set_agents_state first");
    const size_t cnt = (size_t)n * e->ns;
    if (initial)
        for (size_t q = 0; q < cnt; ++q)
            if (!std::isfinite(initial[q])) return fail(e, BFA_ERR_ARG, "internalized value %zu is not finite", q);
    int rc;
    if (!e->trk && (rc = dalloc(e, &e->trk, cnt * 8))) return rc;
    if (e1 || !e->ag_e1) {
        if (!e->ag_e1 && (rc = dalloc(e, &e->ag_e1, cnt * 8))) return rc;
        if (e1)
            HIP_TRY(e, hipMemcpy(e->ag_e1, e1, cnt * 8, hipMemcpyHostToDevice));
        else
            HIP_TRY(e, hipMemset(e->ag_e1, 0, cnt * 8));
    }
    if (initial)
        HIP_TRY(e, hipMemcpy(e->trk, initial, cnt * 8, hipMemcpyHostToDevice));
    else
        HIP_TRY(e, hipMemset(e->trk, 0, cnt * 8));
    return BFA_OK;
}

int bfa_get_internalized(bfa_env* e, double* out) {
    if (!e || !out) return fail(e, BFA_ERR_ARG, "NULL argument");
    if (!e->trk) return fail(e, BFA_ERR_STATE, "internalised-substrate tracking is off");
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    HIP_TRY(e, hipMemcpy(out, e->trk, (size_t)e->n_in * e->ns * 8, hipMemcpyDeviceToHost));
    return BFA_OK;
}

static int ready(bfa_env* e, double dt) { return ensure_precompute(e, dt); }

int bfa_diffusion_decay(bfa_env* e, double dt) {
    if (!e) return BFA_ERR_ARG;
    int rc = ready(e, dt);
    if (rc) return rc;
    return enqueue_diffusion(e);
}

int bfa_secretion_uptake(bfa_env* e) {
    if (!e) return BFA_ERR_ARG;
    return launch_agents(e);
}

int bfa_step(bfa_env* e, double dt, int nsteps) {
    if (!e) return BFA_ERR_ARG;
    if (nsteps < 0) return fail(e, BFA_ERR_ARG, "negative step count");
    int rc = ready(e, dt);
    if (rc) return rc;
    if (nsteps == 0) return BFA_OK;
    if (!e->graph_valid || e->graph_dt != dt) {
        if ((rc = build_graph(e, dt))) return rc;
    }
    for (int s = 0; s < nsteps; ++s) HIP_TRY(e, hipGraphLaunch(e->gexec, e->stream));
    return BFA_OK;
}

int bfa_apply_dirichlet(bfa_env* e) {
    if (!e) return BFA_ERR_ARG;
    // the Dirichlet table lives with the precompute; any dt works for it
    const double dt = e->pre_valid ? e->pre_dt : 1.0;
    int rc = ensure_precompute(e, dt);
    if (rc) return rc;
    return launch_faces(e);
}

int bfa_total_mass(bfa_env* e, double* out) {
    if (!e || !out) return fail(e, BFA_ERR_ARG, "NULL argument");
    std::vector<double> sum(e->ns);
    int rc = reduce(e, e->u, e->prec == BFA_FP32, e->N, 1, e->N, 0, nullptr, sum.data());
    if (rc) return rc;
    const double dV = e->dx * e->dx * e->dx;
    for (int d = 0; d < e->ns; ++d)
        out[d] = (e->prec == BFA_FP32 ? (double)e->N * e->b[d] + e->s[d] * sum[d] : sum[d]) * dV;
    return BFA_OK;
}

int bfa_synchronize(bfa_env* e) {
    if (!e) return BFA_ERR_ARG;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return BFA_OK;
}

// Enqueue the whole-field gradient into g ([ns][3][N], device).
int enqueue_gradient(bfa_env* e, double* g) {
    dim3 grid((unsigned)std::min<long long>(4096, (e->N + 255) / 256), (unsigned)e->ns);
    if (e->prec == BFA_FP32)
        hipLaunchKernelGGL(bfa::k_gradient<float>, grid, dim3(256), 0, e->stream, (const float*)e->u,
                           e->N, e->nx, e->ny, e->nz, e->ns, e->params, e->dx, e->dx, e->dx, g);
    else
        hipLaunchKernelGGL(bfa::k_gradient<double>, grid, dim3(256), 0, e->stream, (const double*)e->u,
                           e->N, e->nx, e->ny, e->nz, e->ns, e->params, e->dx, e->dx, e->dx, g);
    LAUNCH_CHECK(e, "gradient");
    return BFA_OK;
}

int bfa_set_bulk_rates(bfa_env* e, const double* S, const double* T, const double* U, int layout) {
    if (!e) return BFA_ERR_ARG;
    if (!S && !T && !U) {
        dfree(e->bulk_st);
        dfree(e->bulk_ups);
        e->bulk_st = e->bulk_ups = nullptr;
        return BFA_OK;
    }
    if (!S || !T || !U) return fail(e, BFA_ERR_ARG, "S, T and U must all be given (or all NULL)");
    if (layout != BFA_LAYOUT_REF && layout != BFA_LAYOUT_SOA)
        return fail(e, BFA_ERR_ARG, "unknown layout %d", layout);
    const size_t total = (size_t)e->N * e->ns;
    // t2 = T*S and t3 = U + S, each one FP64 rounding, as the reference
    // forms them (BioFVM/BioFVM_microenvironment.cpp:1344,1351)
    std::vector<double> h(2 * total);
    for (size_t a = 0; a < total; ++a) {
        if (!std::isfinite(S[a]) || !std::isfinite(T[a]) || !std::isfinite(U[a]))
            return fail(e, BFA_ERR_ARG, "bulk rate %zu is not finite", a);
        h[a] = T[a] * S[a];
        h[total + a] = U[a] + S[a];
    }
    int rc;
    if (!e->bulk_st && ((rc = dalloc(e, &e->bulk_st, total * 8)) || (rc = dalloc(e, &e->bulk_ups, total * 8)))) {
        dfree(e->bulk_st);
        e->bulk_st = nullptr;
        return rc;
    }
    hipError_t st = hipSuccess;
    if (layout == BFA_LAYOUT_SOA) {
        st = hipMemcpyAsync(e->bulk_st, h.data(), total * 8, hipMemcpyHostToDevice, e->stream);
        if (st == hipSuccess)
            st = hipMemcpyAsync(e->bulk_ups, h.data() + total, total * 8, hipMemcpyHostToDevice, e->stream);
        if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
        if (st != hipSuccess) return fail(e, BFA_ERR_HIP, "bulk rate upload failed: %s", hipGetErrorString(st));
        return BFA_OK;
    }
    double* stage = nullptr;
    double* zo = nullptr;
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(e->stream);
        dfree(stage);
        dfree(zo);
        return code;
    };
    std::vector<double> z1(2 * (size_t)e->ns, 0.0);
    for (int d = 0; d < e->ns; ++d) z1[e->ns + d] = 1.0;
    if ((rc = dalloc(e, &stage, total * 8)) || (rc = dalloc(e, &zo, z1.size() * 8))) return cleanup(rc);
    st = hipMemcpyAsync(zo, z1.data(), z1.size() * 8, hipMemcpyHostToDevice, e->stream);
    dim3 grid((unsigned)(((long long)e->nz * e->ns + 31) / 32), (unsigned)((e->nx + 31) / 32), (unsigned)e->ny);
    for (int f = 0; f < 2 && st == hipSuccess; ++f) {
        st = hipMemcpyAsync(stage, h.data() + f * total, total * 8, hipMemcpyHostToDevice, e->stream);
        if (st != hipSuccess) break;
        hipLaunchKernelGGL(bfa::k_ref_to_soa<double>, grid, dim3(32, 8), 0, e->stream, stage,
                           f == 0 ? e->bulk_st : e->bulk_ups, e->nx, e->ny, e->nz, e->ns, zo, zo + e->ns);
        st = hipGetLastError();
        if (st == hipSuccess) st = hipStreamSynchronize(e->stream);  // stage is reused
    }
    if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "bulk rate upload failed: %s", hipGetErrorString(st)));
    return cleanup(BFA_OK);
}

int bfa_bulk_sources_and_sinks(bfa_env* e, double dt) {
    if (!e) return BFA_ERR_ARG;
    if (!e->bulk_st) return fail(e, BFA_ERR_STATE, "bulk rates not set (bfa_set_bulk_rates)");
    if (!(dt >= 0.0) || !std::isfinite(dt)) return fail(e, BFA_ERR_ARG, "dt must be finite and >= 0");
    dim3 grid((unsigned)std::min<long long>(4096, (e->N + 255) / 256), (unsigned)e->ns);
    if (e->prec == BFA_FP32)
        hipLaunchKernelGGL(bfa::k_bulk<float>, grid, dim3(256), 0, e->stream, (float*)e->u, e->N, e->ns,
                           e->params, e->bulk_st, e->bulk_ups, dt);
    else
        hipLaunchKernelGGL(bfa::k_bulk<double>, grid, dim3(256), 0, e->stream, (double*)e->u, e->N,
                           e->ns, e->params, e->bulk_st, e->bulk_ups, dt);
    LAUNCH_CHECK(e, "bulk sources");
    return BFA_OK;
}

int bfa_gradients(bfa_env* e, double* out, int layout) {
    if (!e || !out) return fail(e, BFA_ERR_ARG, "NULL argument");
    if (layout != BFA_LAYOUT_REF && layout != BFA_LAYOUT_SOA)
        return fail(e, BFA_ERR_ARG, "unknown layout %d", layout);
    const size_t total = (size_t)e->N * e->ns * 3;
    double* g = nullptr;
    double* stage = nullptr;
    double* bs = nullptr;
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(e->stream);
        dfree(g);
        dfree(stage);
        dfree(bs);
        return code;
    };
    int rc = dalloc(e, &g, total * 8);
    if (rc) return rc;
    if ((rc = enqueue_gradient(e, g))) return cleanup(rc);
    const double* src = g;
    if (layout == BFA_LAYOUT_REF) {
        // (substrate, component) pairs are the fastest index of the
        // reference layout: transpose [3ns][N] as a field of 3*ns substrates
        const int n3 = 3 * e->ns;
        std::vector<double> h(2 * (size_t)n3, 0.0);
        for (int q = 0; q < n3; ++q) h[n3 + q] = 1.0;
        if ((rc = dalloc(e, &stage, total * 8)) || (rc = dalloc(e, &bs, h.size() * 8))) return cleanup(rc);
        hipError_t st = hipMemcpyAsync(bs, h.data(), h.size() * 8, hipMemcpyHostToDevice, e->stream);
        if (st == hipSuccess) {
            dim3 grid((unsigned)(((long long)e->nz * n3 + 31) / 32), (unsigned)((e->nx + 31) / 32),
                      (unsigned)e->ny);
            hipLaunchKernelGGL(bfa::k_soa_to_ref<double>, grid, dim3(32, 8), 0, e->stream, g, stage,
                               e->nx, e->ny, e->nz, n3, bs, bs + n3);
            st = hipGetLastError();
        }
        if (st != hipSuccess)
            return cleanup(fail(e, BFA_ERR_HIP, "gradient layout failed: %s", hipGetErrorString(st)));
        src = stage;
    }
    hipError_t st = copy_d2h(out, src, total * 8, e->stream);
    if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "gradient download failed: %s", hipGetErrorString(st)));
    return cleanup(BFA_OK);
}

int bfa_time_gradient(bfa_env* e, int reps, double* ms) {
    if (!e || !ms || reps < 1) return fail(e, BFA_ERR_ARG, "bad argument");
    double* g = nullptr;
    int rc = dalloc(e, &g, (size_t)e->N * e->ns * 3 * 8);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t st = hipEventCreate(&e0);
    if (st == hipSuccess) st = hipEventCreate(&e1);
    if (st == hipSuccess) rc = enqueue_gradient(e, g);  // warm-up
    if (st == hipSuccess && rc == BFA_OK) st = hipEventRecord(e0, e->stream);
    for (int r = 0; r < reps && st == hipSuccess && rc == BFA_OK; ++r) rc = enqueue_gradient(e, g);
    if (st == hipSuccess && rc == BFA_OK) st = hipEventRecord(e1, e->stream);
    if (st == hipSuccess && rc == BFA_OK) st = hipEventSynchronize(e1);
    float t = 0.0f;
    if (st == hipSuccess && rc == BFA_OK) st = hipEventElapsedTime(&t, e0, e1);
    (void)hipStreamSynchronize(e->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(g);
    if (rc) return rc;
    if (st != hipSuccess) return fail(e, BFA_ERR_HIP, "gradient timing failed: %s", hipGetErrorString(st));
    *ms = t / reps;
    return BFA_OK;
}

int bfa_gradients_at(bfa_env* e, int n, const int32_t* voxel, double* out) {
    if (!e || (n > 0 && (!voxel || !out))) return fail(e, BFA_ERR_ARG, "NULL argument");
    if (n < 0) return fail(e, BFA_ERR_ARG, "negative voxel count");
    if (n == 0) return BFA_OK;
    for (int a = 0; a < n; ++a)
        if (voxel[a] < 0 || voxel[a] >= e->N)
            return fail(e, BFA_ERR_ARG, "voxel %d of entry %d outside [0, %lld)", voxel[a], a, e->N);
    int* dv = nullptr;
    double* dout = nullptr;
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(e->stream);
        dfree(dv);
        dfree(dout);
        return code;
    };
    int rc;
    if ((rc = dalloc(e, &dv, (size_t)n * 4)) || (rc = dalloc(e, &dout, (size_t)n * e->ns * 3 * 8)))
        return cleanup(rc);
    hipError_t st = hipMemcpyAsync(dv, voxel, (size_t)n * 4, hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess) {
        const unsigned blocks = (unsigned)(((long long)n * e->ns + 255) / 256);
        if (e->prec == BFA_FP32)
            hipLaunchKernelGGL(bfa::k_gradient_at<float>, dim3(blocks), dim3(256), 0, e->stream,
                               (const float*)e->u, e->N, e->nx, e->ny, e->nz, e->ns, e->params, e->dx,
                               e->dx, e->dx, n, dv, dout);
        else
            hipLaunchKernelGGL(bfa::k_gradient_at<double>, dim3(blocks), dim3(256), 0, e->stream,
                               (const double*)e->u, e->N, e->nx, e->ny, e->nz, e->ns, e->params, e->dx,
                               e->dx, e->dx, n, dv, dout);
        st = hipGetLastError();
    }
    if (st == hipSuccess) st = hipMemcpyAsync(out, dout, (size_t)n * e->ns * 3 * 8, hipMemcpyDeviceToHost, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) return cleanup(fail(e, BFA_ERR_HIP, "gradient download failed: %s", hipGetErrorString(st)));
    return cleanup(BFA_OK);
}

int bfa_time_kernels(bfa_env* e, double dt, int reps, double* ms) {
    if (!e || !ms || reps < 1) return fail(e, BFA_ERR_ARG, "bad argument");
    int rc = ready(e, dt);
    if (rc) return rc;
    const bool slabs = e->slab > 0 && e->R == 0;
    hipEvent_t ev[6];
    for (auto& x : ev) HIP_TRY(e, hipEventCreate(&x));
    double acc[5] = {0, 0, 0, 0, 0};
    for (int r = 0; r < reps && rc == BFA_OK; ++r) {
        (void)hipEventRecord(ev[0], e->stream);
        if (slabs) {
            rc = enqueue_xy_slabs(e);
            (void)hipEventRecord(ev[1], e->stream);
            (void)hipEventRecord(ev[2], e->stream);
            if (rc == BFA_OK) rc = launch_sweep_any(e, 2);
            (void)hipEventRecord(ev[3], e->stream);
        } else {
            for (int a = 0; a < 3 && rc == BFA_OK; ++a) {
                rc = launch_sweep_any(e, a);
                (void)hipEventRecord(ev[a + 1], e->stream);
            }
        }
        if (rc == BFA_OK) rc = launch_agents(e);
        (void)hipEventRecord(ev[4], e->stream);
        (void)hipEventSynchronize(ev[4]);
        float t;
        for (int k = 0; k < 4; ++k) {
            (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
            acc[k] += t;
        }
        (void)hipEventElapsedTime(&t, ev[0], ev[4]);
        acc[4] += t;
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (rc) return rc;
    for (int k = 0; k < 5; ++k) ms[k] = acc[k] / reps;
    return BFA_OK;
}

int bfa_set_solver(bfa_env* e, int variant) {
    if (!e) return BFA_ERR_ARG;
    if (variant != BFA_LOD_3D_BLOCKING && variant != BFA_LOD_3D_AVX256D)
        return fail(e, BFA_ERR_ARG, "unknown solver variant %d", variant);
    if (variant != e->variant) invalidate_graph(e);
    e->variant = variant;
    return BFA_OK;
}

int bfa_get_solver(const bfa_env* e) { return e ? e->variant : -1; }

int bfa_set_dt_caching(bfa_env* e, int keep_first) {
    if (!e) return BFA_ERR_ARG;
    e->keep_first_dt = keep_first != 0;
    return BFA_OK;
}

int bfa_set_slab(bfa_env* e, int planes) {
    if (!e) return BFA_ERR_ARG;
    if (planes < 0) return fail(e, BFA_ERR_ARG, "slab planes must be >= 0, got %d", planes);
    if (planes > 0 && e->R != 0)
        return fail(e, BFA_ERR_ARG, "slab schedule needs the wave-local x sweep (rows_x == 0)");
    e->slab = std::min(planes, e->nz);
    invalidate_graph(e);
    return BFA_OK;
}

int bfa_get_slab(const bfa_env* e) { return e ? e->slab : -1; }

int bfa_set_substrate_streams(bfa_env* e, int groups) {
    if (!e) return BFA_ERR_ARG;
    if (groups != 1 && groups != 2) return fail(e, BFA_ERR_ARG, "substrate stream groups must be 1 or 2, got %d", groups);
    e->split = groups;
    invalidate_graph(e);
    return BFA_OK;
}

int bfa_get_substrate_streams(const bfa_env* e) { return e ? e->split : -1; }

int bfa_set_x_grid(bfa_env* e, int blocks_per_cu) {
    if (!e) return BFA_ERR_ARG;
    if (blocks_per_cu < 0 || blocks_per_cu > 64)
        return fail(e, BFA_ERR_ARG, "x-sweep blocks per CU must be in [0, 64], got %d", blocks_per_cu);
    e->xw_cap = blocks_per_cu;
    invalidate_graph(e);
    return BFA_OK;
}

int bfa_get_x_grid(const bfa_env* e, int* requested, int* resident) {
    if (!e) return BFA_ERR_ARG;
    if (requested) *requested = e->xw_cap;
    if (resident) *resident = e->xw_occ;
    return BFA_OK;
}

int bfa_set_col_runs(bfa_env* e, int runs_y, int runs_z) {
    if (!e) return BFA_ERR_ARG;
    if (runs_y < 1 || runs_y > 64 || runs_z < 1 || runs_z > 64)
        return fail(e, BFA_ERR_ARG, "tile runs per XCD must be in [1, 64], got %d and %d", runs_y, runs_z);
    e->runs[1] = runs_y;
    e->runs[2] = runs_z;
    invalidate_graph(e);
    return BFA_OK;
}

int bfa_get_col_runs(const bfa_env* e, int* runs_y, int* runs_z) {
    if (!e) return BFA_ERR_ARG;
    if (runs_y) *runs_y = e->runs[1];
    if (runs_z) *runs_z = e->runs[2];
    return BFA_OK;
}

int bfa_set_agent_stores(bfa_env* e, int whole_lines) {
    if (!e) return BFA_ERR_ARG;
    if (whole_lines != 0 && whole_lines != 1)
        return fail(e, BFA_ERR_ARG, "agent stores: 1 (whole lines) or 0 (single voxels), got %d", whole_lines);
    if ((whole_lines == 1) != e->agent_lines) invalidate_graph(e);
    e->agent_lines = whole_lines == 1;
    return BFA_OK;
}

int bfa_get_agent_stores(const bfa_env* e) { return e ? (e->agent_lines ? 1 : 0) : -1; }

// Mean milliseconds of `reps` device copies of `bytes` bytes (after one
// untimed launch): the grid-stride k_copy16 or the one-pass k_copy_stream.
static int time_copy(size_t bytes, int reps, double* ms, bool stream) {
    if (!ms || reps < 1 || bytes < 16) return BFA_ERR_ARG;
    const long long n = (long long)(bytes / 16);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return BFA_ERR_NOMEM;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return BFA_ERR_NOMEM;
    }
    (void)hipMemset(a, 1, n * 16);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const unsigned blocks = stream ? (unsigned)((n + 1023) / 1024)
                                   : (unsigned)std::min<long long>(65536, (n + 255) / 256);
    auto launch = [&] {
        if (stream)
            hipLaunchKernelGGL(bfa::k_copy_stream, dim3(blocks), dim3(256), 0, nullptr,
                               reinterpret_cast<const bfa::u32x4*>(a), reinterpret_cast<bfa::u32x4*>(b), n);
        else
            hipLaunchKernelGGL(bfa::k_copy16, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
    };
    launch();
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, nullptr);
    hipError_t st = hipEventSynchronize(e1);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (st != hipSuccess) return BFA_ERR_HIP;
    *ms = t / reps;
    return BFA_OK;
}

int bfa_copy_kernel(size_t bytes, int reps, double* ms) { return time_copy(bytes, reps, ms, false); }

int bfa_copy_kernel_stream(size_t bytes, int reps, double* ms) {
    if (bytes > (size_t)1 << 40) return BFA_ERR_ARG;  // grid.x stays below 2^31
    return time_copy(bytes, reps, ms, true);
}

int bfa_set_tiles(bfa_env* e, int Lx, int Ly, int Lz, int rows_x, int cols_y, int cols_z) {
    if (!e) return BFA_ERR_ARG;
    int sLreq[3], sL[3], sTX[3], sTXreq[3];
    std::copy(e->Lreq, e->Lreq + 3, sLreq);
    std::copy(e->L, e->L + 3, sL);
    std::copy(e->TX, e->TX + 3, sTX);
    std::copy(e->TXreq, e->TXreq + 3, sTXreq);
    const int sRreq = e->Rreq, sR = e->R;
    e->Lreq[0] = Lx;
    e->Lreq[1] = Ly;
    e->Lreq[2] = Lz;
    e->Rreq = rows_x;
    e->TXreq[1] = cols_y;
    e->TXreq[2] = cols_z;
    int rc = apply_tiles(e);
    if (rc) {
        std::copy(sLreq, sLreq + 3, e->Lreq);
        std::copy(sL, sL + 3, e->L);
        std::copy(sTX, sTX + 3, e->TX);
        std::copy(sTXreq, sTXreq + 3, e->TXreq);
        e->Rreq = sRreq;
        e->R = sR;
        return rc;
    }
    if (e->R != 0) e->slab = 0;
    invalidate_graph(e);
    return BFA_OK;
}

int bfa_get_tiles(const bfa_env* e, int* out6) {
    if (!e || !out6) return BFA_ERR_ARG;
    out6[0] = e->L[0];
    out6[1] = e->L[1];
    out6[2] = e->L[2];
    out6[3] = e->R;
    out6[4] = e->TX[1];
    out6[5] = e->TX[2];
    return BFA_OK;
}

}  // extern "C"

// ---- batches of independent microenvironments ------------------------------
#include "biofvm_batch_host.inc"
#include "biofvm_agents_dev.inc"
#include "biofvm_noalloc.inc"
