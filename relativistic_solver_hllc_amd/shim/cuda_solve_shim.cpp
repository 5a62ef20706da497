// Drop-in implementation of the reference's GPU offload host API
// (solver_hllc_illum_cuda/cuda/cuda_solve.h:5-37, cuda/cuda_solve.cu,
// cuda/cuda_struct.cu) over the C ABI of librsh_scatter.so.
//
// Build inside the reference tree (replacing cuda/*.cu in makefile:35,163-173):
//   g++ -std=c++17 -DRSH_SHIM_REFERENCE_TYPES -include solve_module/solve_global_struct.h \
//       -include global_value.h -I<repo>/include -c cuda_solve_shim.cpp && link with -lrsh_scatter
// Build stand-alone (tests): g++ -std=c++17 -I<repo>/include cuda_solve_shim.cpp ...
#include <cstdio>
#include <cstdlib>
#include <string>
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
    bool geometry = false;  // normals/squares/volume uploaded (start == 0 and files present)
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

// Reads one simple-bin geometry file (int count + raw elements of elem_size
// bytes) and checks its count; false if it is missing or has the wrong size.
bool read_geometry_file(const std::string& path, int32_t elem_size, int64_t expect, std::vector<double>& out) {
    int64_t count = 0;
    if (rsh_read_simple_bin(path.c_str(), elem_size, nullptr, 0, &count) != RSH_OK) {
        std::printf("Warning: %s\n", rsh_last_error());
        return false;
    }
    if (count != expect) {
        std::printf("Warning: %s holds %lld elements, expected %lld\n", path.c_str(), (long long)count,
                    (long long)expect);
        return false;
    }
    out.resize((size_t)count * (elem_size / sizeof(double)));
    return rsh_read_simple_bin(path.c_str(), elem_size, out.data(), count, &count) == RSH_OK;
}

// normals.bin / squares.bin / volume.bin of glb_files.base_adress
// (cuda/cuda_struct.cu:137-147; ReadNormalFile reader_bin.cpp:49-69 reads
// `base` Vector3 per cell, ReadSimpleFileBin reader_bin.h:13-31 the others).
void load_geometry(int64_t n, int faces) {
    std::vector<double> normals, areas, volume;
    const std::string base = glb_files.base_adress;
    const bool ok = read_geometry_file(base + "normals.bin", faces * 3 * (int32_t)sizeof(double), n, normals) &&
                    read_geometry_file(base + "squares.bin", (int32_t)sizeof(double), n * faces, areas) &&
                    read_geometry_file(base + "volume.bin", (int32_t)sizeof(double), n, volume);
    if (!ok) {
        std::printf("Warning: no cell geometry; divstream/divimpuls unavailable\n");
        return;
    }
    check(rsh_problem_set_geometry(g.problem, normals.data(), areas.data(), volume.data()), "InitDevice(geometry)");
    g.geometry = true;
}

// Launches the moments on stream s with the requested host copies (pinned
// grid arrays); sync waits for them, as the reference's blocking copies do.
int moments(grid_t& grid, int s, bool energy, bool stream, bool impuls, bool divstream, bool divimpuls, bool sync,
            const char* where) {
    require_init(where);
    if ((divstream || divimpuls) && !g.geometry) {
        std::printf("Error: %s needs normals.bin/squares.bin/volume.bin (InitDevice with start == 0)\n", where);
        std::exit(1);
    }
    if (!grid.energy || !grid.divstream) {
        std::printf("Error: %s: moment arrays exist only on the rank with start == 0\n", where);
        std::exit(1);
    }
    check(rsh_moments(g.problem, s, energy ? as_doubles(grid.energy) : nullptr,
                      stream ? as_doubles(grid.stream) : nullptr, impuls ? as_doubles(grid.impuls) : nullptr,
                      divstream ? as_doubles(grid.divstream) : nullptr, divimpuls ? as_doubles(grid.divimpuls) : nullptr),
          where);
    if (sync) check(rsh_sync(g.problem, s), where);
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
    g.geometry = false;
    if (start == 0) {  // only the first rank computes the moments (cuda_struct.cu:61,158)
        const int64_t n = grid_host.size;
        const int widths[5] = {1, 3, 1, 3, 9};  // divstream, divimpuls, energy, stream, impuls
        void* arr[5] = {};
        for (int i = 0; i < 5; i++)
            check(rsh_host_alloc_pinned(widths[i] * n * (int64_t)sizeof(double), &arr[i]), "InitDevice");
        grid_host.divstream = static_cast<Type*>(arr[0]);
        grid_host.divimpuls = reinterpret_cast<decltype(grid_host.divimpuls)>(arr[1]);
        grid_host.energy = static_cast<Type*>(arr[2]);
        grid_host.stream = reinterpret_cast<decltype(grid_host.stream)>(arr[3]);
        grid_host.impuls = reinterpret_cast<decltype(grid_host.impuls)>(arr[4]);
        load_geometry(n, d.faces);
    }
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
    // upload, integral and copy back overlapped over cell ranges; returns when done
    check(rsh_scatter_from_host(g.problem, grid.Illum, g.k_begin, g.k_end, grid.scattering, 32),
          "CalculateIntScattering");
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

// Radiation moments (cuda/cuda_solve.cu:172-320). The reference's
// CalculateEnergy / CalculateDivStream / CalculateDivImpuls have their kernel
// launches commented out and only copy device arrays back; here each computes
// what its name says. All use stream eCuda_params and return after the copies.
int CalculateEnergy(const grid_directions_t&, grid_t& grid) {
    return moments(grid, eCuda_params, true, false, false, false, false, true, "CalculateEnergy");
}
int CalculateStream(const grid_directions_t&, grid_t& grid) {
    return moments(grid, eCuda_params, false, true, false, false, false, true, "CalculateStream");
}
int CalculateImpuls(const grid_directions_t&, grid_t& grid) {
    return moments(grid, eCuda_params, false, false, true, false, false, true, "CalculateImpuls");
}
// With CUDA_FULL_ARRAYS the reference also copies impuls / stream back (cuda_solve.cu:242-244,265-267).
int CalculateDivImpuls(const grid_directions_t&, grid_t& grid) {
    return moments(grid, eCuda_params, false, false, true, false, true, true, "CalculateDivImpuls");
}
int CalculateDivStream(const grid_directions_t&, grid_t& grid) {
    return moments(grid, eCuda_params, false, true, false, true, false, true, "CalculateDivStream");
}
// cuda_solve.cu:272-295: kernel, blocking copies of divstream/divimpuls, and
// (CUDA_FULL_ARRAYS) energy/stream/impuls.
int CalculateAllParam(const grid_directions_t&, grid_t& grid) {
    return moments(grid, eCuda_params, true, true, true, true, true, true, "CalculateAllParam");
}
// cuda_solve.cu:297-320: the same on stream st, asynchronously; the caller
// waits with CudaSyncStream(st) (Illum_async.cpp:692 and later).
int CalculateAllParamStream(const grid_directions_t&, grid_t& grid, e_cuda_stream_id_t st) {
    return moments(grid, (int)st, true, true, true, true, true, false, "CalculateAllParamStream");
}
