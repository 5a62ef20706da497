// Replays the call sequence of MPI_CalculateIllumAsync (solver_hllc_illum_cuda/
// solve_module/illum/Illum_async.cpp:663-1072) on one rank through the drop-in
// names of rsh_cuda_solve.hpp:
//   SetDevice, InitDevice(start, end), CudaSendIllumAsync (block of the first
//   section, then direction by direction with the 4-argument overload),
//   CalculateIntScatteringAsyncMPIStream(0, size_1, stream 1) and
//   (0, local, stream 2), CalculateAllParamStream(eCuda_params) on the rank
//   with start == 0 (Illum_async.cpp:692), CudaSyncStream, CudaEventPrint,
//   CalculateIntScattering, CalculateEnergy, ClearDevice, ClearHost (grid_t
//   destructor).
// Usage: shim_driver <input> <output> [<geometry dir/>]; the geometry dir holds
// normals.bin, squares.bin and volume.bin (glb_files.base_adress).
// Input file (little-endian): int64 N, int32 M, int32 start, int32 end,
// double xyz[M*3], double w[M], double F, double illum[4*M*N].
// Output file: double S_async[(end-start)*N], double S_sync[(end-start)*N] and,
// when start == 0, energy[N], stream[N*3], impuls[N*9], divstream[N],
// divimpuls[N*3] from CalculateAllParamStream, then energy[N] from CalculateEnergy.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rsh_cuda_solve.hpp"

global_files_t glb_files;  // the reference defines it in main.cpp:6

int main(int argc, char** argv) {
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: shim_driver <input> <output> [<geometry dir/>]\n");
        return 2;
    }
    if (argc == 4) glb_files.base_adress = argv[3];
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t N = 0;
    int32_t M = 0, start = 0, end = 0;
    bool ok = std::fread(&N, 8, 1, f) == 1 && std::fread(&M, 4, 1, f) == 1 && std::fread(&start, 4, 1, f) == 1 &&
              std::fread(&end, 4, 1, f) == 1;
    std::vector<double> xyz(3 * (size_t)M), w(M), illum(4 * (size_t)M * N);
    double F = 0;
    ok = ok && std::fread(xyz.data(), 8, xyz.size(), f) == xyz.size() && std::fread(w.data(), 8, w.size(), f) == w.size() &&
         std::fread(&F, 8, 1, f) == 1 && std::fread(illum.data(), 8, illum.size(), f) == illum.size();
    std::fclose(f);
    if (!ok) return 2;

    grid_directions_t dirs;
    dirs.size = M;
    dirs.directions.resize(M);
    for (int i = 0; i < M; i++) {
        dirs.directions[i].dir = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        dirs.directions[i].area = w[i];
    }
    dirs.full_area = F;

    std::vector<double> s_async, s_sync, moments, energy2;
    {
        grid_t grid;
        grid.size = (int)N;
        SetDevice(0);
        InitDevice(dirs, grid, start, end);
        for (size_t i = 0; i < illum.size(); i++) grid.Illum[i] = illum[i];

        const int local = end - start;
        const int size_1 = local / 3;  // size_1_section = send_count[0] / 3 (Illum_async.cpp:104)
        const int per_dir = 4 * (int)N;
        // section 1 of every direction block in one message, then the rest direction by direction
        CudaSendIllumAsync(size_1 * per_dir, 0, grid.Illum);
        for (int j = size_1; j < M; j++) CudaSendIllumAsync(per_dir, j * per_dir, j * per_dir, grid.Illum);

        CalculateIntScatteringAsyncMPIStream(dirs, grid, 0, size_1, eCuda_scattering_1);
        CalculateIntScatteringAsyncMPIStream(dirs, grid, 0, local, eCuda_scattering_2);
        if (start == 0) CalculateAllParamStream(dirs, grid, eCuda_params);
        CudaSyncStream(eCuda_scattering_1);
        CudaSyncStream(eCuda_scattering_2);
        if (start == 0) {
            CudaSyncStream(eCuda_params);
            const double* arrays[5] = {grid.energy, grid.stream, grid.impuls, grid.divstream, grid.divimpuls};
            const int widths[5] = {1, 3, 9, 1, 3};
            for (int i = 0; i < 5; i++) moments.insert(moments.end(), arrays[i], arrays[i] + widths[i] * N);
        }
        CudaEventPrint();
        s_async.assign(grid.scattering, grid.scattering + (size_t)local * N);

        for (size_t i = 0; i < (size_t)local * N; i++) grid.scattering[i] = -1.0;
        CalculateIntScattering(dirs, grid);
        CudaWait();
        s_sync.assign(grid.scattering, grid.scattering + (size_t)local * N);
        if (start == 0) {
            for (int64_t i = 0; i < N; i++) grid.energy[i] = -1.0;
            CalculateEnergy(dirs, grid);
            energy2.assign(grid.energy, grid.energy + N);
        }
        ClearDevice();
    }  // ~grid_t -> ClearHost

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    std::fwrite(s_async.data(), 8, s_async.size(), o);
    std::fwrite(s_sync.data(), 8, s_sync.size(), o);
    std::fwrite(moments.data(), 8, moments.size(), o);
    std::fwrite(energy2.data(), 8, energy2.size(), o);
    std::fclose(o);
    std::printf("shim_driver ok N=%lld M=%d range=[%d,%d)\n", (long long)N, M, start, end);
    return 0;
}
