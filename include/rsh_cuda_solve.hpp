// rsh_cuda_solve.hpp -- C++ drop-in for the reference's GPU offload header
// solver_hllc_illum_cuda/cuda/cuda_solve.h:5-37 (Lartsev-Artem/relativistic_solver_hllc).
//
// Declares the reference's host functions with their exact names and
// signatures, implemented in relativistic_solver_hllc_amd/shim/cuda_solve_shim.cpp
// on top of the C ABI in rsh_scatter.h (librsh_scatter.so). Semantics kept:
// module-global device state, library-owned pinned host buffers grid.Illum /
// grid.scattering, exit(1) with a message on any failure (cuda/cuda_def.cuh:14-20).
// Differences (all fixes of reference defects):
//   - the D2H copy of CalculateIntScatteringAsyncMPIStream lands at
//     grid.scattering + start*N (the reference writes + start, cuda_solve.cu:107);
//   - uploads are ordered before later scattering launches on every stream;
//   - 64-bit overloads of CudaSendIllumAsync for offsets beyond 2^31;
//   - the three streams keep the default priority.
//
// Radiation moments (CalculateAllParam*, CalculateEnergy, ...) run the
// library's moments kernel. InitDevice with start == 0 reads the cell geometry
// from glb_files.base_adress + normals.bin / squares.bin / volume.bin as the
// reference does (cuda/cuda_struct.cu:137-147) and allocates the pinned host
// moment arrays grid.divstream / divimpuls / energy / stream / impuls (the
// reference's ON_FULL_ILLUM_ARRAYS set). Missing geometry files only disable
// the divergences, which then exit(1) with a message when requested.
//
// Types: with RSH_SHIM_REFERENCE_TYPES defined, include the reference's
// solve_module/solve_global_struct.h and global_value.h before this header and
// its grid_t, grid_directions_t, e_cuda_stream_id_t, Type and glb_files are
// used. Otherwise a minimal layout-compatible set is defined here (same member
// names; dir[] indexable) and the program defines `global_files_t glb_files;`
// (as the reference's main.cpp:6 does).
#pragma once

#include <cstdint>

#ifndef RSH_SHIM_REFERENCE_TYPES
#include <array>
#include <string>
#include <vector>

typedef double Type;

// solve_module/solve_global_struct.h:234-240
enum e_cuda_stream_id_t { eCuda_scattering_1 = 0, eCuda_scattering_2, eCuda_params, eCuda_count };

// global_def.h:79-90
struct direction_s {
    std::array<Type, 3> dir;
    Type area;
};
struct grid_directions_t {
    int size = 0;
    std::vector<direction_s> directions;
    Type full_area = 0;
};

// global_value.h:97-188: only the member the GPU offload reads.
struct global_files_t {
    std::string base_adress;
};
extern global_files_t glb_files;

void ClearHost(struct grid_t& grid_host);

// The fields of grid_t (solve_global_struct.h:153-230) that the GPU offload uses.
struct grid_t {
    int size = 0;
    Type* Illum = nullptr;
    Type* scattering = nullptr;
    Type* divstream = nullptr;
    Type* divimpuls = nullptr;   // Vector3 per cell in the reference (3 doubles)
    Type* energy = nullptr;
    Type* stream = nullptr;      // Vector3 per cell
    Type* impuls = nullptr;      // Matrix3 per cell (9 doubles)
    ~grid_t() { ClearHost(*this); }  // solve_global_struct.cpp:135-143
};
#endif

void CudaEventPrint();
void CudaSyncStream(const e_cuda_stream_id_t stream_id);
int CalculateIntScatteringAsyncMPIStream(const grid_directions_t& grid_dir, grid_t& grid, const int start,
                                         const int end, const e_cuda_stream_id_t stream);

void CudaSendIllumAsync(const int size, const int shift, const Type* Illum_host);
void CudaSendIllumAsync(const int size, const int shift_dev, const int shift_host, const Type* Illum_host);
void CudaSendIllumAsync(const int64_t size, const int64_t shift, const Type* Illum_host);
void CudaSendIllumAsync(const int64_t size, const int64_t shift_dev, const int64_t shift_host,
                        const Type* Illum_host);
int CalculateIntScatteringAsyncMPI(const grid_directions_t& grid_dir, grid_t& grid, int local_size);
void CopyIllumOnDevice(const int size, const Type* Illum_host);

int CalculateIntScattering(const grid_directions_t& grid_dir, grid_t& grid);
int CalculateIntScatteringAsync(const grid_directions_t& grid_dir, grid_t& grid);

int CalculateEnergy(const grid_directions_t& grid_dir, grid_t& grid);
int CalculateStream(const grid_directions_t& grid_dir, grid_t& grid);
int CalculateImpuls(const grid_directions_t& grid_dir, grid_t& grid);
int CalculateDivImpuls(const grid_directions_t& grid_dir, grid_t& grid);
int CalculateDivStream(const grid_directions_t& grid_dir, grid_t& grid);
int CalculateAllParam(const grid_directions_t& grid_dir, grid_t& grid);
int CalculateAllParamStream(const grid_directions_t& grid_dir, grid_t& grid, e_cuda_stream_id_t st);

void SetDevice(const int num_dev);
void InitDevice(const grid_directions_t& grid_dir_host, grid_t& grid_host, const int start, const int end);
void ClearDevice();

void ClearHost(grid_t& grid_host);
void CudaWait();
