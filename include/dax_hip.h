/* dax_hip.h -- C-ABI of the gfx950 (MI355X) device adapter for Dax's
 * marching-cubes isosurface path.
 *
 * Plain pointers and sizes only.  Every pointer named d_* is device memory;
 * h_* is host memory.  `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Every call enqueues on `stream` and returns without waiting,
 * except the ones documented as synchronous (they read a result back).
 * Every function returns a status: DAX_HIP_OK (0) or one of the error codes
 * below, with a message in dax_hip_last_error() (thread-local).
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the Kitware/DaxToolkit source tree).
 */
#ifndef DAX_HIP_H
#define DAX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Replace the exception types of the reference:
 * dax/cont/ErrorControlOutOfMemory.h, dax/cont/ErrorExecution.h,
 * dax/cont/ErrorControlBadValue.h (thrown at e.g.
 * dax/cuda/cont/internal/ArrayManagerExecutionCuda.h:46-84,
 * dax/thrust/cont/internal/DeviceAdapterAlgorithmThrust.h:409-434). */
#define DAX_HIP_OK 0
#define DAX_HIP_OOM 1
#define DAX_HIP_EXEC_ERROR 2
#define DAX_HIP_BAD_VALUE 3
#define DAX_HIP_CAPACITY 4 /* outputs larger than the buffers given */

/* Adapter id, next to DAX_DEVICE_ADAPTER_{ERROR,UNDEFINED,SERIAL,CUDA,OPENMP,TBB}
 * (dax/cont/internal/DeviceAdapterTag.h:22-41). */
#define DAX_DEVICE_ADAPTER_HIP 5

/* ---- runtime ---------------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none). */
const char* dax_hip_last_error(void);

/* Library version string, e.g. "daxhip 0.1 gfx950". */
const char* dax_hip_version(void);

/* Number of visible HIP devices; synchronous. */
int dax_hip_device_count(int* h_count);

/* ---- memory: ArrayManagerExecution<T,C,Hip> ---------------------------
 * Replaces ArrayManagerExecutionThrustDevice's device_vector storage
 * (dax/thrust/cont/internal/ArrayManagerExecutionThrustDevice.h:65-186) and
 * the CUDA manager's OOM handling (ArrayManagerExecutionCuda.h:33-85):
 * allocation failure returns DAX_HIP_OOM with the sticky error cleared. */
int dax_hip_malloc(void** d_ptr, size_t bytes);
int dax_hip_free(void* d_ptr);
/* LoadDataForInput / RetrieveOutputData / CopyInto transfers. */
int dax_hip_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, void* stream);
int dax_hip_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, void* stream);
int dax_hip_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream);
int dax_hip_memset(void* d_dst, int value, size_t bytes, void* stream);
/* DeviceAdapterAlgorithm<Tag>::Synchronize
 * (dax/cuda/cont/internal/DeviceAdapterAlgorithmCuda.h:40-43). */
int dax_hip_synchronize(void* stream);

/* ---- timer: DeviceAdapterTimerImplementation<Hip> ---------------------
 * Replaces the CUDA event timer (DeviceAdapterAlgorithmCuda.h:49-88). */
int dax_hip_timer_create(void** h_timer);
int dax_hip_timer_destroy(void* timer);
int dax_hip_timer_reset(void* timer, void* stream);
/* Seconds since the last reset; synchronises on the stop event. */
int dax_hip_timer_elapsed(void* timer, void* stream, double* h_seconds);

/* ---- marching cubes on a uniform grid -----------------------------------
 * Grid: point dims (nx, ny, nz), flat point index i + nx*(j + ny*k)
 * (dax/Extent.h:104-112), coordinates origin + spacing*ijk
 * (dax/cont/UniformGrid.h:82-89).  Cells (nx-1)(ny-1)(nz-1), x fastest.
 *
 * The forward runs as three kernels over a caller-provided workspace:
 *   classify: field -> one inside-bit per point (f > iso, strict, as
 *             GetHexahedronClassification, dax/worklet/MarchingCubes.h:36-48)
 *   count+scan: per 64-point row word, active-edge and triangle counts, and a
 *             single-pass decoupled look-back exclusive scan
 *             (replaces ScanInclusive at DispatcherGenerateInterpolatedCells.h:180-183)
 *   generate: vertices, normals, connectivity (MarchingCubesGenerate,
 *             MarchingCubes.h:90-169, plus ResolveCoordinates
 *             DispatcherGenerateInterpolatedCells.h:261-301)
 * With dedup on (the reference default, :67,77) vertex u is the u-th active
 * edge in (lower point id, axis x<y<z) order, which equals the reference's
 * sort/unique/lower-bounds numbering; with dedup off the 3T triangle corners
 * are output in triangle order with identity connectivity. */

/* Workspace bytes needed for a volume of nx*ny*nz points. */
size_t dax_hip_mc_workspace_bytes(int64_t nx, int64_t ny, int64_t nz);

/* Flags for dax_hip_mc_forward. */
#define DAX_HIP_MC_DEDUP 1u          /* RemoveDuplicatePoints(true) */
#define DAX_HIP_MC_NORMALS 2u        /* write d_normals */

/* classify + count/scan + generate in one call, graph-capturable (no host
 * read-back).  d_counts (device int64[2]) receives {T, U}: triangles and
 * output points.  Outputs are written only if T <= cap_tris and
 * U <= cap_points; otherwise nothing is written and the counts still tell the
 * sizes needed (check with dax_hip_mc_read_counts).  Optional outputs may be
 * NULL: d_normals [U][3] f32, d_edges [U] u32 compact active-edge list
 * ((axis << 30) | lower point id; dedup only; needs P <= 2^30),
 * d_interp_first/second [U] i32 and d_interp_weight [U] f32 (the
 * reference's InterpolationWeights, EdgeInterpolatedGrid.h:28-74, kept for
 * CompactPointField).  d_vertices [U][3] f32 and d_conn [3T] i32 are
 * required. */
int dax_hip_mc_forward(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                       const float* h_origin3, const float* h_spacing3, float iso,
                       unsigned flags, void* d_workspace, size_t workspace_bytes,
                       float* d_vertices, float* d_normals, int32_t* d_conn,
                       uint32_t* d_edges, int32_t* d_interp_first,
                       int32_t* d_interp_second, float* d_interp_weight,
                       int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                       void* stream);

/* Synchronous: reads {T, U} written by a forward (or classify) on `stream`,
 * and returns DAX_HIP_CAPACITY if the generate step was skipped for lack of
 * room, DAX_HIP_EXEC_ERROR if T or U left the int32 dax::Id range. */
int dax_hip_mc_read_counts(const int64_t* d_counts, int64_t cap_points, int64_t cap_tris,
                           int64_t* h_tris, int64_t* h_points, void* stream);

/* The three kernels of dax_hip_mc_forward as separate calls on one
 * workspace, in this order (each enqueues exactly one kernel):
 *   dax_hip_mc_bits      classify: field -> inside-bits (also resets the
 *                        scan state of the workspace)
 *   dax_hip_mc_scan      per-word counts + decoupled look-back scan
 *   dax_hip_mc_generate  outputs, as dax_hip_mc_forward (d_counts written)
 * Together they replace DispatcherMapCell<MarchingCubesCount>
 * (dax/cont/DispatcherMapCell.h:27-64), the ScanInclusive/UpperBounds of
 * DispatcherGenerateInterpolatedCells.h:180-223 and its generate/resolve
 * steps (:239-301). */
int dax_hip_mc_bits(const float* d_field, int64_t nx, int64_t ny, int64_t nz, float iso,
                    void* d_workspace, size_t workspace_bytes, void* stream);
int dax_hip_mc_scan(int64_t nx, int64_t ny, int64_t nz, void* d_workspace,
                    size_t workspace_bytes, void* stream);
int dax_hip_mc_generate(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                        const float* h_origin3, const float* h_spacing3, float iso,
                        unsigned flags, const void* d_workspace, size_t workspace_bytes,
                        float* d_vertices, float* d_normals, int32_t* d_conn,
                        uint32_t* d_edges, int32_t* d_interp_first,
                        int32_t* d_interp_second, float* d_interp_weight,
                        int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                        void* stream);

/* First two steps only (classify + count/scan); {T, U_dedup} to d_counts.
 * Replaces DispatcherMapCell<MarchingCubesCount> + the inclusive scan. */
int dax_hip_mc_classify(const float* d_field, int64_t nx, int64_t ny, int64_t nz, float iso,
                        void* d_workspace, size_t workspace_bytes, int64_t* d_counts,
                        void* stream);

/* After dax_hip_mc_classify on the same workspace: the per-cell outputs of
 * the reference's count step.  Any pointer may be NULL.
 *   d_cases   [C] u8   case index (GetHexahedronClassification)
 *   d_count   [C] i32  NumFaces[case] (MarchingCubesCount's FieldOut)
 *   d_offsets [C] i32  exclusive scan of d_count (first triangle id of each
 *                      cell; ScanExclusive semantics,
 *                      DeviceAdapterAlgorithmSerial.h:70-99) */
int dax_hip_mc_cell_outputs(int64_t nx, int64_t ny, int64_t nz, const void* d_workspace,
                            size_t workspace_bytes, uint8_t* d_cases, int32_t* d_count,
                            int32_t* d_offsets, void* stream);

/* ---- tile-shape tuning --------------------------------------------------
 * params4 = {classify_unroll, classify_vec, scan_sub, gen_tile}:
 *   classify_vec 1 (one float per lane, unroll 8/16/32 row words per wave)
 *   or 4 (float4 per lane, unroll 2/4/8 chunks of 256 points; rows of a
 *   multiple of 64 points); scan_sub 1/2/4/8 (scan tile = scan_sub*1024
 *   row words); gen_tile 256/512/1024 (row words per generate block).
 * A shape of (0,0,0) sets or gets the default for shapes not tuned. */
int dax_hip_mc_set_tuning(int64_t nx, int64_t ny, int64_t nz, const int* h_params4);
int dax_hip_mc_get_tuning(int64_t nx, int64_t ny, int64_t nz, int* h_params4);
/* Times every variant of each kernel with HIP events (reps launches each),
 * keeps the fastest per kernel for this shape and returns it (h_best4) with
 * the per-kernel milliseconds (h_best_ms3, may be NULL).  Synchronous.  The
 * outputs and workspace are those of dax_hip_mc_forward. */
int dax_hip_mc_autotune(const float* d_field, int64_t nx, int64_t ny, int64_t nz, float iso,
                        unsigned flags, void* d_workspace, size_t workspace_bytes,
                        float* d_vertices, float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                        int64_t cap_points, int64_t cap_tris, int64_t* d_counts, int reps,
                        int* h_best4, float* h_best_ms3, void* stream);
/* Persist / restore the per-shape choices (text, one shape per line). */
int dax_hip_mc_tuning_save(const char* path);
int dax_hip_mc_tuning_load(const char* path);

/* Interpolate any scalar point field onto the output points with the
 * interpolation records of a forward: out[u] = Lerp(in[first], in[second], w)
 * (CompactPointField, DispatcherGenerateInterpolatedCells.h:107-141;
 * InterpolateFieldToField, dax/exec/internal/kernel/GenerateWorklets.h:65-100). */
int dax_hip_mc_interpolate_scalar(const float* d_in, const int32_t* d_first,
                                  const int32_t* d_second, const float* d_weight,
                                  int64_t n, float* d_out, void* stream);

/* ---- adjoint (no reference counterpart) ---------------------------------
 * dL/dfield of L(vertices, normals) given dL/dvertices [U][3] and
 * dL/dnormals [U][3], with the topology of the forward held fixed.  Uses the
 * forward's workspace (inside-bits and vertex bases must still hold the
 * forward of this field and iso) and its compact edge list d_edges; the
 * interpolants (t, endpoint gradients) are recomputed from the field.
 * U = min(d_counts[1], n_points) if d_counts is given (device {T, U} of the
 * forward), else n_points.  d_scratch needs
 * dax_hip_mc_backward_scratch_bytes(n_points) bytes.  Deterministic: no
 * atomics, fixed summation order.  d_dl_dfield [P] f32 is fully written. */
size_t dax_hip_mc_backward_scratch_bytes(int64_t n_points);
int dax_hip_mc_backward(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                        const float* h_origin3, const float* h_spacing3, float iso,
                        const void* d_workspace, size_t workspace_bytes, const uint32_t* d_edges,
                        const int64_t* d_counts, int64_t n_points, const float* d_dl_dvertices,
                        const float* d_dl_dnormals, void* d_scratch, size_t scratch_bytes,
                        float* d_dl_dfield, void* stream);

/* Central-difference point gradient of the whole field, [P][3] f32 (interior
 * (f(p+e)-f(p-e))/(2h), one-sided on the first/last plane).  No reference
 * counterpart; the normals of dax_hip_mc_forward use the same stencil. */
int dax_hip_point_gradient(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                           const float* h_spacing3, float* d_grad, void* stream);

/* CellGradient on a uniform grid, [C][3] f32 (dax/worklet/CellGradient.h:28-48,
 * CellDerivativeAxisAligned dax/exec/Derivative.h:37-75 at the voxel centre). */
int dax_hip_cell_gradient(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                          const float* h_origin3, const float* h_spacing3, float* d_grad,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DAX_HIP_H */
