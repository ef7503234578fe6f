// fvm_kernels.hpp -- host-side launch interface of the kernels in fvm_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include "fvm_device.hpp"

namespace fvm {

enum KernelOp { OP_STAGE1, OP_STAGE1_FUSED, OP_STAGE2, OP_DIFFUSE, OP_DIFFUSE_FINAL, OP_BC_INPLACE, OP_MAXIMA };

struct KernelCall {
    KernelOp op;
    dim3 grid;
    hipStream_t stream;
    const void* cases;
    const WorkItem* work;
    const int* count;
    Roles ri, rv;
    bool kahan;
    // non-null: start / stop events recorded on the dispatch (timing mode)
    hipEvent_t ev0, ev1;
};

bool launch_tiled(bool fp64, int tx, int ty, const KernelCall& kc);
bool tile_supported(int tx, int ty, bool fp64);
// Smallest tile width and height among the compiled shapes.
void tile_min_dims(int* tx, int* ty);
// Workgroups per CU resident at once for the stage kernels a step launches
// (stage 1, stage 2, the fused stage 1 when `fused`) and for the two half
// steps, from the occupancy calculator: the persistent grids launch that many
// per CU, so no workgroup waits for another to finish its whole item list.
bool tiled_resident_per_cu(bool fp64, int tx, int ty, bool kahan, bool fused, int* stage, int* diffuse);
// Whether stage 1 with the first viscous half step fused in is built for a tile.
bool fused_d1_supported(int tx, int ty, bool fp64);
// Tiles of one item of the per-tile-maxima reduction: case_list holds one
// item (case, chunk) per REDUCE_CHUNK tiles of every active case.
constexpr int REDUCE_CHUNK = 1024;

// Ghost fill over a list of ghost cells.  final_: the step's last launch,
// which also takes the ghost cells' CFL maxima and, in workgroups
// [gblocks, grid.x), reduces the per-tile maxima of the items in case_list.
// single_case >= 0: the list is (single_case, 0 .. n-1) and is not read.
// dt64: the engine keeps its CFL maxima in FP64 (FP64 state, compensated variant).
void launch_ghost(bool fp64, bool dt64, bool final_, dim3 grid, hipStream_t s, const void* c, const WorkItem* items,
                  const int* n, Roles ri, Roles rv, int gblocks = 0, const WorkItem* case_list = nullptr,
                  const int* ncases = nullptr, int single_case = -1, hipEvent_t ev0 = nullptr,
                  hipEvent_t ev1 = nullptr);
void launch_ghost_prep(bool fp64, hipStream_t s, const void* c, const WorkItem* items, int n);
void launch_mark(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int fixup, int fix_set);
void launch_init(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int set_a, int set_b);
void launch_copy_set(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int src, int dst);
void launch_zero_u32(hipStream_t s, unsigned* p, int n);
void launch_field(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int set, int field, float* out);
// Velocity arrows of set `set` over the nax x nay sample cells (every step-th
// cell) into out, at most max_vertices vertices; *counter (zeroed by the
// caller) ends as the vertices the arrows need.
void launch_arrows(bool fp64, hipStream_t s, const void* c, int case_id, int set, int step, int nax, int nay,
                   float u_inf, float max_len, float cos_a, float sin_a, float head_ratio, int max_vertices, float* out,
                   unsigned* counter);
void launch_stat(bool fp64, hipStream_t s, const void* c, int case_id, size_t ncells, int set, int mode, unsigned* target);
// FP64 totals of four planes of n cells (FP32 or FP64 values): nblocks
// partials of 4 doubles in `partial`, then their fixed-order sum into out[4].
// Every caller takes kConsBlocks workgroups, so the totals of a state are the
// same bits whichever entry point computes them.
constexpr int kConsBlocks = 1024;
void launch_cons(bool fp64, hipStream_t s, const void* const planes[4], size_t n, double* partial, int nblocks,
                 double* out);
void launch_forces(bool fp64, hipStream_t s, const void* c, int case_id, int set, double* partial, int nblocks, double* out);

}  // namespace fvm
