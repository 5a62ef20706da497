// Drop-in implementation of the reference's GPU offload host API
// (solver_hllc_illum_cuda/cuda/cuda_solve.h:5-37, cuda/cuda_solve.cu,
// cuda/cuda_struct.cu) over the C ABI of librsh_scatter.so.
//
// Build inside the reference tree (replacing cuda/*.cu in makefile:35,163-173):
//   g++ -std=c++17 -DRSH_SHIM_REFERENCE_TYPES -include solve_module/solve_global_struct.h \
//       -I<repo>/include -c cuda_solve_shim.cpp && link with -lrsh_scatter
// Build stand-alone (tests): g++ -std=c++17 -I<repo>/include cuda_solve_shim.cpp ...
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rsh_cuda_solve.hpp"
#include "rsh_scatter.h"

namespace {

// Module-global state, as the reference's device_host_ptr / grid_*_device_ptr
// (cuda/cuda_struct.cu:5-7).
struct ShimState {
    int device = 0;
    rsh_problem problem = nullptr;
    int k_begin = 0, k_end = 0;
    int64_t n_cells = 0;
    int n_dirs = 0;
};
ShimState g;

// Any failure prints and exits, as CUDA_CALL_FUNC does (cuda/cuda_def.cuh:18-20).
void check(int status, const char* where) {
    if (status != RSH_OK) {
        std::printf("Error rsh code: %d \n%s\nCalls from: %s\n", status, rsh_last_error(), where);
        std::exit(1);
    }
}

void require_init(const char* where) {
    if (!g.problem) {
        std::printf("Error: %s called before InitDevice\n", where);
        std::exit(1);
    }
}

template <class P>
double* as_doubles(P* p) {
    return reinterpret_cast<double*>(p);
}

// Launch local output rows [start, end) on stream id s and copy them to
// grid.scattering + start * N (reference: cuda_solve.cu:89-111).
int scatter_local(grid_t& grid, int start, int end, int s) {
    require_init("CalculateIntScattering*");
    if (start < 0 || end > g.k_end - g.k_begin || start > end) {
        std::printf("Error: local direction range [%d,%d) outside [0,%d)\n", start, end, g.k_end - g.k_begin);
        std::exit(1);
    }
    check(rsh_scatter(g.problem, g.k_begin + start, g.k_begin + end, s, grid.scattering), "rsh_scatter");
    return 0;
}

}  // namespace

void SetDevice(const int num_dev) {
    check(rsh_set_device(num_dev), "SetDevice");
    g.device = num_dev;
}

// Reference: cuda/cuda_struct.cu:134-178,187-190. start/end are the global
// output directions of this rank (solve_main.cpp:136).
void InitDevice(const grid_directions_t& grid_dir_host, grid_t& grid_host, const int start, const int end) {
    if (g.problem) {
        std::printf("Error: InitDevice called twice without ClearDevice\n");
        std::exit(1);
    }
    const int m = grid_dir_host.size;
    std::vector<double> xyz(3 * (size_t)m), w(m);
    for (int i = 0; i < m; i++) {
        xyz[3 * i] = grid_dir_host.directions[i].dir[0];
        xyz[3 * i + 1] = grid_dir_host.directions[i].dir[1];
        xyz[3 * i + 2] = grid_dir_host.directions[i].dir[2];
        w[i] = grid_dir_host.directions[i].area;
    }
    rsh_problem_desc d{};
    d.device = g.device;
    d.n_cells = grid_host.size;
    d.faces = 4;  // base = NUMBER_OF_MEASUREMENTS + 1 (global_def.h:23, prj_config.h:18)
    d.n_dirs = m;
    d.dirs_xyz = xyz.data();
    d.weights = w.data();
    d.full_area = grid_dir_host.full_area;
    d.k_begin = start;
    d.k_end = end;
    d.phase = RSH_PHASE_THOMSON;
    d.g = 0.0;
    d.normalize = 0;
    check(rsh_create(&d, &g.problem), "InitDevice");
    g.k_begin = start;
    g.k_end = end;
    g.n_cells = grid_host.size;
    g.n_dirs = m;
    // Pinned host arrays owned by the module (cuda_struct.cu:154-156).
    void* p = nullptr;
    check(rsh_host_alloc_pinned((int64_t)4 * m * grid_host.size * (int64_t)sizeof(double), &p), "InitDevice");
    grid_host.Illum = static_cast<Type*>(p);
    check(rsh_host_alloc_pinned((int64_t)(end - start) * grid_host.size * (int64_t)sizeof(double), &p),
          "InitDevice");
    grid_host.scattering = static_cast<Type*>(p);
}

void ClearDevice() {
    if (g.problem) check(rsh_destroy(g.problem), "ClearDevice");
    g.problem = nullptr;
}

// Reference: cuda/cuda_struct.cu:197-214 (called from grid_t::~grid_t).
void ClearHost(grid_t& grid_host) {
    check(rsh_host_free_pinned(grid_host.Illum), "ClearHost");
    check(rsh_host_free_pinned(grid_host.scattering), "ClearHost");
    check(rsh_host_free_pinned(as_doubles(grid_host.divstream)), "ClearHost");
    check(rsh_host_free_pinned(as_doubles(grid_host.divimpuls)), "ClearHost");
    check(rsh_host_free_pinned(as_doubles(grid_host.energy)), "ClearHost");
    check(rsh_host_free_pinned(as_doubles(grid_host.stream)), "ClearHost");
    check(rsh_host_free_pinned(as_doubles(grid_host.impuls)), "ClearHost");
    grid_host.Illum = nullptr;
    grid_host.scattering = nullptr;
    grid_host.divstream = nullptr;
    grid_host.divimpuls = nullptr;
    grid_host.energy = nullptr;
    grid_host.stream = nullptr;
    grid_host.impuls = nullptr;
}

// Reference: cuda/cuda_solve.cu:69-83 (always stream 0 = eCuda_scattering_1).
void CudaSendIllumAsync(const int size, const int shift, const Type* Illum_host) {
    CudaSendIllumAsync((int64_t)size, (int64_t)shift, (int64_t)shift, Illum_host);
}
void CudaSendIllumAsync(const int size, const int shift_dev, const int shift_host, const Type* Illum_host) {
    CudaSendIllumAsync((int64_t)size, (int64_t)shift_dev, (int64_t)shift_host, Illum_host);
}
void CudaSendIllumAsync(const int64_t size, const int64_t shift, const Type* Illum_host) {
    CudaSendIllumAsync(size, shift, shift, Illum_host);
}
void CudaSendIllumAsync(const int64_t size, const int64_t shift_dev, const int64_t shift_host,
                        const Type* Illum_host) {
    require_init("CudaSendIllumAsync");
    check(rsh_upload_illum(g.problem, Illum_host, size, shift_dev, shift_host, RSH_STREAM_SCATTERING_1),
          "CudaSendIllumAsync");
}

// Reference: cuda/cuda_solve.cu:85-88.
void CopyIllumOnDevice(const int size, const Type* Illum_host) {
    require_init("CopyIllumOnDevice");
    check(rsh_copy_illum_sync(g.problem, Illum_host, size), "CopyIllumOnDevice");
}

int CalculateIntScatteringAsyncMPIStream(const grid_directions_t&, grid_t& grid, const int start, const int end,
                                         const e_cuda_stream_id_t stream) {
    return scatter_local(grid, start, end, (int)stream);
}

// Reference: cuda/cuda_solve.cu:112-130 (default stream -> stream 0 here).
int CalculateIntScatteringAsyncMPI(const grid_directions_t&, grid_t& grid, int local_size) {
    return scatter_local(grid, 0, local_size, RSH_STREAM_SCATTERING_1);
}

// Reference: cuda/cuda_solve.cu:132-151: every local output direction.
int CalculateIntScatteringAsync(const grid_directions_t&, grid_t& grid) {
    return scatter_local(grid, 0, g.k_end - g.k_begin, RSH_STREAM_SCATTERING_1);
}

// Reference: cuda/cuda_solve.cu:153-170: upload all of grid.Illum, compute, copy back, synchronously.
int CalculateIntScattering(const grid_directions_t&, grid_t& grid) {
    require_init("CalculateIntScattering");
    check(rsh_copy_illum_sync(g.problem, grid.Illum, (int64_t)4 * g.n_dirs * g.n_cells), "CalculateIntScattering");
    scatter_local(grid, 0, g.k_end - g.k_begin, RSH_STREAM_SCATTERING_1);
    check(rsh_sync(g.problem, RSH_STREAM_SCATTERING_1), "CalculateIntScattering");
    return 0;
}

// Reference: cuda/cuda_solve.cu:37-45.
void CudaSyncStream(const e_cuda_stream_id_t stream_id) {
    require_init("CudaSyncStream");
    check(rsh_sync(g.problem, (int)stream_id), "CudaSyncStream");
}

void CudaWait() {
    if (g.problem) check(rsh_sync(g.problem, -1), "CudaWait");
    else check(rsh_dev_synchronize(g.device), "CudaWait");
}

// Reference: cuda/cuda_solve.cu:16-36 reads events it never creates; here the
// last scattering kernel time of each stream is printed.
void CudaEventPrint() {
    require_init("CudaEventPrint");
    float ms[RSH_STREAM_COUNT];
    check(rsh_event_times(g.problem, ms), "CudaEventPrint");
    std::printf("time kernel1 %f\n", ms[RSH_STREAM_SCATTERING_1]);
    std::printf("time kernel2 %f\n\n", ms[RSH_STREAM_SCATTERING_2]);
}

// Radiation moments (cuda/cuda_solve.cu:172-320) are outside the scattering
// path this library implements; they report an error instead of returning
// unfilled arrays.
static int moments_unavailable(const char* name) {
    std::printf("Error: %s (radiation moments) is not provided by librsh_scatter\n", name);
    return 1;
}
int CalculateEnergy(const grid_directions_t&, grid_t&) { return moments_unavailable("CalculateEnergy"); }
int CalculateStream(const grid_directions_t&, grid_t&) { return moments_unavailable("CalculateStream"); }
int CalculateImpuls(const grid_directions_t&, grid_t&) { return moments_unavailable("CalculateImpuls"); }
int CalculateDivImpuls(const grid_directions_t&, grid_t&) { return moments_unavailable("CalculateDivImpuls"); }
int CalculateDivStream(const grid_directions_t&, grid_t&) { return moments_unavailable("CalculateDivStream"); }
int CalculateAllParam(const grid_directions_t&, grid_t&) { return moments_unavailable("CalculateAllParam"); }
int CalculateAllParamStream(const grid_directions_t&, grid_t&, e_cuda_stream_id_t) {
    return moments_unavailable("CalculateAllParamStream");
}
