// Internal types of the gfx950 solver library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sfp {

// ------------------------------------------------------------------ errors
struct HipError : std::runtime_error {
    explicit HipError(const std::string &m) : std::runtime_error(m) {}
};
struct ArgError : std::runtime_error {
    explicit ArgError(const std::string &m) : std::runtime_error(m) {}
};
struct InputError : std::runtime_error {
    explicit InputError(const std::string &m) : std::runtime_error(m) {}
};
struct StateError : std::runtime_error {
    explicit StateError(const std::string &m) : std::runtime_error(m) {}
};

void hip_fail(hipError_t e, const char *what, const char *file, int line);
// Sets the text sfp_last_error(NULL) returns (entry points without a context).
void set_unbound_error(const char *msg);
#define SFP_HIP(call)                                                    \
    do {                                                                 \
        hipError_t _e = (call);                                          \
        if (_e != hipSuccess) ::sfp::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)
#define SFP_LAUNCH_CHECK() SFP_HIP(hipGetLastError())

// ------------------------------------------------------------------ grid
// Cell-grid extents; linear index = z + nz*(y + ny*x) (core/Grid.h:28).
struct Dims {
    int nx = 0, ny = 0, nz = 0;
    __host__ __device__ size_t cells() const { return (size_t)nx * ny * nz; }
    __host__ __device__ size_t idx(int i, int j, int k) const {
        return (size_t)k + (size_t)nz * ((size_t)j + (size_t)ny * (size_t)i);
    }
    // extents of face grid a (core/StaggeredGrid.h:15-19)
    __host__ __device__ Dims face(int a) const {
        Dims f = *this;
        if (a == 0) f.nx += 1;
        if (a == 1) f.ny += 1;
        if (a == 2) f.nz += 1;
        return f;
    }
    __host__ __device__ Dims coarse() const { return Dims{(nx + 1) / 2, (ny + 1) / 2, (nz + 1) / 2}; }
    __host__ __device__ int n(int a) const { return a == 0 ? nx : (a == 1 ? ny : nz); }
};

// Launch shape for a cell-parallel kernel: a 1-D grid of 256-thread blocks,
// each covering a tz x ty tile (z, the unit-stride axis, across the 64 lanes
// of a wavefront; y across wavefronts) of `planes` consecutive x planes.
struct CellGrid {
    unsigned gx, gy;  // tiles per x plane along z and y
    int planes;       // consecutive x planes per block
};
struct Launch {
    dim3 grid, block;
    CellGrid cg;
    unsigned blocks() const { return grid.x; }
};
// planes <= 0 picks the number of x planes per block from the grid size.
// planes <= 0: choose up to max_planes planes per block, keeping >= ~2048 blocks
Launch cell_launch(Dims d, int planes = 0, int max_planes = 16);

// ------------------------------------------------------------------ device memory
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void ensure(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        SFP_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// ------------------------------------------------------------------ operators
// A 7-point symmetric operator on a cell grid:
//   (A x)_c = diag_c x_c - sum_a ( cpl_a[c] x_{c-e_a} + cpl_a[c+e_a] x_{c+e_a} )
// cpl_a[c] >= 0 couples c - e_a and c (0 on the lower boundary, at inactive
// cells and across solid faces). Inactive cells are identity rows (diag 1)
// with zero right-hand side and zero solution.
struct Op64 {
    const double *diag;
    const double *cpl[3];
    // 1: each diagonal entry equals the sum of the row's six couplings, added
    // in the order -x, +x, -y, +y, -z, +z (0 on the boundary), bit for bit,
    // except at cell 0 (an identity row) and its upper neighbours 1, nz and
    // nz*ny (whose coupling to cell 0 is eliminated but stays in the
    // diagonal). Kernels may then rebuild the diagonal from the couplings
    // they load anyway instead of reading it.
    int pinned_rowsum = 0;
};
struct Op32 {
    const float *diag;  // diagonal for active cells, 0 marks an inactive cell
    const float *cpl[3];
    // 1: as Op64::pinned_rowsum, with the FP32 sums (cell 0's rebuilt
    // diagonal is 0, an inactive cell, as stored; cells 1, nz and nz*ny differ)
    int pinned_rowsum = 0;
};

// One multigrid level: FP64 operator (setup, and the fine level's outer
// Krylov solve) and its FP32 copy (V-cycle).
struct Level {
    Dims d;
    DevBuf<double> diag64, cpl64[3];
    DevBuf<float> diag32, cpl32[3];
    DevBuf<float> x0, x1, dv, b;  // ping-pong iterates, Chebyshev direction, rhs
    // z rows (i, j) holding an active cell (levels whose step pairs run over
    // whole rows; null elsewhere and where every row is active)
    DevBuf<unsigned char> rowact;
    // both operators have the row-sum structure of Op64::pinned_rowsum and
    // Op32::pinned_rowsum, and kernels may rebuild the diagonals
    bool pinned_rowsum = false;
    Op64 op64() const {
        return Op64{diag64.p, {cpl64[0].p, cpl64[1].p, cpl64[2].p}, pinned_rowsum ? 1 : 0};
    }
    Op32 op32() const {
        return Op32{diag32.p, {cpl32[0].p, cpl32[1].p, cpl32[2].p}, pinned_rowsum ? 1 : 0};
    }
};

// Scalars of the flexible-CG recurrences, resident on the device.
struct KrylovScalars {
    double zw, zq, zr, pq_prev;  // dots of the current iteration
    double alpha, beta, pq;
    double rr, bb;               // ||r||^2, ||b||^2
    double pad[4];               // scratch of single reductions (count, scales)
    // batched iterations: `stop` becomes nonzero when rr <= stop_rr after an
    // update, and every gated kernel of later iterations then returns at once;
    // `iters` counts the updates done
    double stop, iters, stop_rr;
    double alpha_prev;  // the previous iteration's alpha (lagged solution update)
    // learned preconditioner in a batch: an update whose ||r||^2 exceeds
    // stall_rr times the previous one raises `stalled` and `stop` (0: off)
    double stall_rr, stalled;
};

// Per-kernel timing collected with HIP events on the context stream.
enum KernelId {
    K_SMOOTH = 0,   // fine-level Chebyshev step reading and writing the direction (FP32)
    K_MATVEC64,     // FCG matvec + dots (FP64)
    K_UPDATE64,     // FCG vector update (FP64)
    K_RESTRICT,     // fine residual + restriction
    K_SMOOTH_PAIR,  // fine-level last two post-smoothing steps in one pass (FP32)
    K_SMOOTH_RESTRICT,  // fine-level last pre-smoothing step + restriction in one pass
    K_COUNT
};

}  // namespace sfp
