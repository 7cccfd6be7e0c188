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
 * allocation failure returns DAX_HIP_OOM with the sticky error cleared.
 * Blocks come from the device's memory pool in null-stream order (a freed
 * block is reused once the work queued before the free has finished); work
 * on other streams that uses a block must finish before it is freed. */
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

/* ---- DeviceAdapterAlgorithm<DeviceAdapterTagHip> ----------------------- */

/* DispatcherMapField<Magnitude> (dax/cont/DispatcherMapField.h:30-61 with
 * dax/worklet/Magnitude.h:26-33): out[i] = Sqrt(dot(in[i], in[i])) over n
 * packed float32 triples, the dot product summed from component 0. */
int dax_hip_magnitude_vec3(const float* d_in, int64_t n, float* d_out, void* stream);

/* The total ScanInclusive returns (DeviceAdapterAlgorithm.h:114-117;
 * DispatcherGenerateInterpolatedCells.h:180-183 takes T from it) without
 * writing the scan: *h_total = sum of n int32 values, in 64 bits.
 * Synchronous. */
int dax_hip_reduce_sum_i32(const int32_t* d_in, int64_t n, int64_t* h_total, void* stream);

/* Copy (DeviceAdapterAlgorithm.h:56-60; Thrust CopyPortal
 * DeviceAdapterAlgorithmThrust.h:76-83): a device copy kernel. */
int dax_hip_copy(void* d_dst, const void* d_src, size_t bytes, void* stream);

/* ScanInclusive / ScanExclusive (DeviceAdapterAlgorithm.h:114-130; Serial
 * DeviceAdapterAlgorithmSerial.h:46-99): d_out may equal d_in; *h_total
 * (may be NULL) receives the full sum, as both reference scans return it.
 * Synchronous.  Deterministic summation order (reduce-then-scan). */
int dax_hip_scan_inclusive_i32(const int32_t* d_in, int32_t* d_out, int64_t n, int32_t* h_total, void* stream);
int dax_hip_scan_exclusive_i32(const int32_t* d_in, int32_t* d_out, int64_t n, int32_t* h_total, void* stream);
int dax_hip_scan_inclusive_i64(const int64_t* d_in, int64_t* d_out, int64_t n, int64_t* h_total, void* stream);
int dax_hip_scan_exclusive_i64(const int64_t* d_in, int64_t* d_out, int64_t n, int64_t* h_total, void* stream);
int dax_hip_scan_inclusive_f32(const float* d_in, float* d_out, int64_t n, float* h_total, void* stream);
int dax_hip_scan_exclusive_f32(const float* d_in, float* d_out, int64_t n, float* h_total, void* stream);
int dax_hip_scan_inclusive_f64(const double* d_in, double* d_out, int64_t n, double* h_total, void* stream);
int dax_hip_scan_exclusive_f64(const double* d_in, double* d_out, int64_t n, double* h_total, void* stream);

/* LowerBounds / UpperBounds (DeviceAdapterAlgorithm.h:62-78, 297-317;
 * General :258-308, :1044-1092): for each value, the first index of the
 * sorted array whose element is not less (lower) / is greater (upper) than
 * it; `descending` != 0 for arrays sorted with SortGreater.  Output as int32
 * (d_out32) and/or int64 (d_out64). */
int dax_hip_lower_bounds_i32(const int32_t* d_sorted, int64_t n, const int32_t* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);
int dax_hip_upper_bounds_i32(const int32_t* d_sorted, int64_t n, const int32_t* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);
int dax_hip_lower_bounds_i64(const int64_t* d_sorted, int64_t n, const int64_t* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);
int dax_hip_upper_bounds_i64(const int64_t* d_sorted, int64_t n, const int64_t* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);
int dax_hip_lower_bounds_f32(const float* d_sorted, int64_t n, const float* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);
int dax_hip_upper_bounds_f32(const float* d_sorted, int64_t n, const float* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);
int dax_hip_lower_bounds_f64(const double* d_sorted, int64_t n, const double* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);
int dax_hip_upper_bounds_f64(const double* d_sorted, int64_t n, const double* d_values, int64_t m, int descending, int32_t* d_out32, int64_t* d_out64, void* stream);

/* Sort / SortByKey / Unique (DeviceAdapterAlgorithm.h:183-250, 271-281):
 * stable LSD radix sort, ascending (SortLess) or descending (SortGreater);
 * SortByKey permutes value records of value_bytes bytes with the keys;
 * Unique removes consecutive duplicates in place and returns the new size.
 * Synchronous. */
int dax_hip_sort_i32(int32_t* d_values, int64_t n, int descending, void* stream);
int dax_hip_sort_i64(int64_t* d_values, int64_t n, int descending, void* stream);
int dax_hip_sort_u32(uint32_t* d_values, int64_t n, int descending, void* stream);
int dax_hip_sort_u64(uint64_t* d_values, int64_t n, int descending, void* stream);
int dax_hip_sort_f32(float* d_values, int64_t n, int descending, void* stream);
int dax_hip_sort_f64(double* d_values, int64_t n, int descending, void* stream);
int dax_hip_sort_by_key_i32(int32_t* d_keys, void* d_values, size_t value_bytes, int64_t n, int descending, void* stream);
int dax_hip_sort_by_key_i64(int64_t* d_keys, void* d_values, size_t value_bytes, int64_t n, int descending, void* stream);
int dax_hip_sort_by_key_u32(uint32_t* d_keys, void* d_values, size_t value_bytes, int64_t n, int descending, void* stream);
int dax_hip_sort_by_key_u64(uint64_t* d_keys, void* d_values, size_t value_bytes, int64_t n, int descending, void* stream);
int dax_hip_sort_by_key_f32(float* d_keys, void* d_values, size_t value_bytes, int64_t n, int descending, void* stream);
int dax_hip_sort_by_key_f64(double* d_keys, void* d_values, size_t value_bytes, int64_t n, int descending, void* stream);
int dax_hip_unique_i32(int32_t* d_values, int64_t n, int64_t* h_new_n, void* stream);
int dax_hip_unique_i64(int64_t* d_values, int64_t n, int64_t* h_new_n, void* stream);
int dax_hip_unique_u32(uint32_t* d_values, int64_t n, int64_t* h_new_n, void* stream);
int dax_hip_unique_u64(uint64_t* d_values, int64_t n, int64_t* h_new_n, void* stream);
int dax_hip_unique_f32(float* d_values, int64_t n, int64_t* h_new_n, void* stream);
int dax_hip_unique_f64(double* d_values, int64_t n, int64_t* h_new_n, void* stream);

/* StreamCompact (DeviceAdapterAlgorithm.h:252-269; General :769-830):
 * keeps the entries whose stencil element is non-zero.  With d_in == NULL
 * the kept indices are written (d_indices32 and/or d_indices64, the
 * reference's StreamCompact(stencil, output)); otherwise the kept records of
 * value_bytes bytes go to d_out.  stencil_type: 0 int32, 1 int64, 2 float,
 * 3 uint8, 4 double.  *h_count = number kept.  Synchronous. */
int dax_hip_stream_compact(const void* d_in, size_t value_bytes, const void* d_stencil,
                           int stencil_type, int64_t n, void* d_out, int32_t* d_indices32,
                           int64_t* d_indices64, int64_t* h_count, void* stream);

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
#define DAX_HIP_TOPOLOGY_MASK_OPTIONAL 4u /* topology prepare: the int32 point mask may be
                                            left unwritten (CompactPointField still works) */

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
/* dax_hip_mc_forward that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops).  The workspace is a 48 MiB array the
 * library holds on each GPU it runs on (grids up to 512^3 points; larger
 * ones return DAX_HIP_BAD_VALUE); every call on a GPU uses that GPU's array,
 * so calls that overlap on one GPU race.  Arguments as dax_hip_mc_forward
 * without the workspace. */
int dax_hip_mc_forward_noalloc(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                               const float* h_origin3, const float* h_spacing3, float iso,
                               unsigned flags, float* d_vertices, float* d_normals,
                               int32_t* d_conn, uint32_t* d_edges, int32_t* d_interp_first,
                               int32_t* d_interp_second, float* d_interp_weight,
                               int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                               void* stream);

/* Slice of a uniform grid by a plane: replaces DispatcherMapCell<SliceCount>
 * + DispatcherGenerateInterpolatedCells<SliceGenerate> over the grid's
 * point coordinates (dax/worklet/Slice.h:28-173; the reference test
 * dax/worklet/testing/UnitTestWorkletSlice.cxx:66-121).  The classified and
 * interpolated value of a point is dot(plane_normal, c - plane_origin) at
 * its coordinates c = origin + spacing * ijk, against the isovalue
 * dot(plane_normal, plane_origin) -- the reference's definition -- computed
 * in the kernels from the point's indices: no field is read.  Outputs,
 * flags, capacities and counts as dax_hip_mc_forward (normals are the
 * central-difference gradients of those values; the reference's slice has
 * none).  dax_hip_mc_slice_classify is the counting step alone (as
 * dax_hip_mc_classify). */
int dax_hip_mc_slice_classify(int64_t nx, int64_t ny, int64_t nz, const float* h_origin3,
                              const float* h_spacing3, const float* h_plane_origin3,
                              const float* h_plane_normal3, void* d_workspace,
                              size_t workspace_bytes, int64_t* d_counts, void* stream);
int dax_hip_mc_slice(int64_t nx, int64_t ny, int64_t nz, const float* h_origin3,
                     const float* h_spacing3, const float* h_plane_origin3,
                     const float* h_plane_normal3, unsigned flags, void* d_workspace,
                     size_t workspace_bytes, float* d_vertices, float* d_normals, int32_t* d_conn,
                     uint32_t* d_edges, int32_t* d_interp_first, int32_t* d_interp_second,
                     float* d_interp_weight, int64_t cap_points, int64_t cap_tris,
                     int64_t* d_counts, void* stream);

/* Marching tetrahedra on an unstructured tetrahedral grid: replaces
 * DispatcherMapCell<MarchingTetrahedraCount> +
 * DispatcherGenerateInterpolatedCells<MarchingTetrahedraGenerate>
 * (dax/worklet/MarchingTetrahedra.h:29-143; the reference test
 * dax/worklet/testing/UnitTestWorkletMarchingTetrahedra.cxx:59-135).
 * d_tets [n_tets][4] int32 point ids, d_field [n_points] f32, d_points
 * [n_points][3] f32 coordinates.
 * dax_hip_mt_count (synchronous): d_counts [n_tets] = NumFaces[case],
 * d_offsets [n_tets] their exclusive scan, *h_tris = T; an id outside
 * [0, n_points) is DAX_HIP_BAD_VALUE.
 * dax_hip_mt_generate (synchronous with DAX_HIP_MC_DEDUP): the triangles'
 * corners as interpolation records (first < second point id, weight
 * (iso - f_first) / (f_second - f_first)) and coordinates (Lerp of the
 * endpoints' coordinates); without dedup 3T points in corner order and
 * identity connectivity (d_first/second/weight required), with dedup the
 * unique edges sorted by (first, second) and connectivity by lower bound
 * (the reference's Sort/Unique/LowerBounds, :261-301), d_first/second/
 * weight optional.  d_vertices [3T][3] and d_conn [3T] (3T is the most
 * points there can be); *h_points = U.  Dedup needs
 * dax_hip_mt_scratch_bytes(T) bytes of d_scratch. */
int dax_hip_mt_count(const int32_t* d_tets, int64_t n_tets, const float* d_field,
                     int64_t n_points, float iso, int32_t* d_counts, int32_t* d_offsets,
                     int64_t* h_tris, void* stream);
size_t dax_hip_mt_scratch_bytes(int64_t n_tris);
int dax_hip_mt_generate(const int32_t* d_tets, int64_t n_tets, const float* d_field,
                        const float* d_points, int64_t n_points, float iso, unsigned flags,
                        const int32_t* d_counts, const int32_t* d_offsets, int64_t n_tris,
                        void* d_scratch, size_t scratch_bytes, float* d_vertices,
                        int32_t* d_conn, int32_t* d_first, int32_t* d_second, float* d_weight,
                        int64_t* h_points, void* stream);

/* Marching cubes on an unstructured hexahedral grid: replaces
 * DispatcherMapCell<MarchingCubesCount> +
 * DispatcherGenerateInterpolatedCells<MarchingCubesGenerate> on an
 * UnstructuredGrid<CellTagHexahedron> (dax/worklet/MarchingCubes.h:53-164;
 * the reference test runs on it through TryAllGridTypes,
 * dax/worklet/testing/UnitTestWorkletMarchingCubes.cxx:211-216).
 * d_hexes [n_hexes][8] int32 point ids in the hexahedron's vertex order
 * (dax/cont/testing/TestingGridGenerator.h:232-241).  The same contract as
 * dax_hip_mt_count / dax_hip_mt_generate above, with the marching-cubes
 * case table and the hexahedron's 12 edges (MarchingCubes.h:131-135).
 * No normals: the reference computes none on this path. */
int dax_hip_mc_hex_count(const int32_t* d_hexes, int64_t n_hexes, const float* d_field,
                         int64_t n_points, float iso, int32_t* d_counts, int32_t* d_offsets,
                         int64_t* h_tris, void* stream);
size_t dax_hip_mc_hex_scratch_bytes(int64_t n_tris);
int dax_hip_mc_hex_generate(const int32_t* d_hexes, int64_t n_hexes, const float* d_field,
                            const float* d_points, int64_t n_points, float iso, unsigned flags,
                            const int32_t* d_counts, const int32_t* d_offsets, int64_t n_tris,
                            void* d_scratch, size_t scratch_bytes, float* d_vertices,
                            int32_t* d_conn, int32_t* d_first, int32_t* d_second,
                            float* d_weight, int64_t* h_points, void* stream);

/* Threshold (dax/worklet/Threshold.h:30-110) through DispatcherMapCell and
 * DispatcherGenerateTopology (dax/cont/DispatcherGenerateTopology.h:146-256).
 *
 * dax_hip_threshold_count_uniform / _cells replace
 * DispatcherMapCell<ThresholdCount<Scalar>>: d_counts[c] = 1 when every
 * point value v of cell c has min <= v <= max, else 0.  Uniform grids: cells
 * x-fastest, (nx-1)(ny-1)(nz-1) of them, d_field [nx*ny*nz].  Unstructured:
 * d_cells [n_cells][nv] int32 point ids (1 <= nv <= 8); an id outside
 * [0, n_points) is DAX_HIP_BAD_VALUE (synchronous).
 *
 * dax_hip_topology_prepare_* / _generate_* replace
 * DispatcherGenerateTopology<ThresholdTopology> for any non-negative counts
 * (a cell is repeated count times, in cell order).  prepare (synchronous):
 * *h_cells = T output cells; with DAX_HIP_MC_DEDUP (RemoveDuplicatePoints,
 * the reference's default) d_point_mask [n_points] int32 = 1 where a kept
 * cell uses the point (the reference's PointMask, for CompactPointField =
 * dax_hip_stream_compact with it as the stencil) and *h_points = U used
 * points.  generate (after prepare with the same counts and workspace):
 * d_conn [T][nv] = the input cells' vertex ids (uniform grids: the voxel's 8
 * corners in the hexahedron's order), remapped with dedup to the points'
 * ranks among the used points (the reference's Sort/Unique/LowerBounds),
 * and d_coords [U][3] the used points' coordinates in point order (uniform:
 * origin + spacing * ijk); without dedup the ids are the input's and no
 * coordinates are written.  Workspace: dax_hip_topology_workspace_bytes. */
int dax_hip_threshold_count_uniform(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                                    float min_value, float max_value, int32_t* d_counts,
                                    void* stream);
int dax_hip_threshold_count_cells(const int32_t* d_cells, int nv, int64_t n_cells,
                                  const float* d_field, int64_t n_points, float min_value,
                                  float max_value, int32_t* d_counts, void* stream);
size_t dax_hip_topology_workspace_bytes(int64_t n_cells, int64_t n_points);
/* ThresholdCount and the prepare step in one call on a uniform grid (the
 * reference's Threshold pipeline, Benchmarks/Threshold/Pipeline.h:163-190):
 * d_counts as dax_hip_threshold_count_uniform, then as
 * dax_hip_topology_prepare_uniform on them.  With rows of 16-byte aligned
 * float4s (nx % 4 == 0) the count kernel also forms the bit-word layout's
 * cell words.  With DAX_HIP_TOPOLOGY_MASK_OPTIONAL, d_point_mask may be
 * left unwritten where that layout is used; pass the buffer anyway. */
int dax_hip_threshold_prepare_uniform(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                                      float min_value, float max_value, unsigned flags,
                                      int32_t* d_counts, void* d_workspace,
                                      size_t workspace_bytes, int32_t* d_point_mask,
                                      int64_t* h_cells, int64_t* h_points, void* stream);
int dax_hip_topology_prepare_uniform(int64_t nx, int64_t ny, int64_t nz, const int32_t* d_counts,
                                     unsigned flags, void* d_workspace, size_t workspace_bytes,
                                     int32_t* d_point_mask, int64_t* h_cells, int64_t* h_points,
                                     void* stream);
int dax_hip_topology_generate_uniform(int64_t nx, int64_t ny, int64_t nz, const float* h_origin3,
                                      const float* h_spacing3, const int32_t* d_counts,
                                      unsigned flags, const void* d_workspace,
                                      size_t workspace_bytes, const int32_t* d_point_mask,
                                      int32_t* d_conn, float* d_coords, void* stream);
int dax_hip_topology_prepare_cells(const int32_t* d_cells, int nv, int64_t n_cells,
                                   int64_t n_points, const int32_t* d_counts, unsigned flags,
                                   void* d_workspace, size_t workspace_bytes,
                                   int32_t* d_point_mask, int64_t* h_cells, int64_t* h_points,
                                   void* stream);
int dax_hip_topology_generate_cells(const int32_t* d_cells, int nv, int64_t n_cells,
                                    const float* d_points, int64_t n_points,
                                    const int32_t* d_counts, unsigned flags,
                                    const void* d_workspace, size_t workspace_bytes,
                                    const int32_t* d_point_mask, int32_t* d_conn,
                                    float* d_coords, void* stream);
/* CompactPointField after a prepare with DAX_HIP_MC_DEDUP (same workspace and
 * mask): d_out[U] = the used points' values of d_in [n_points] (value_bytes
 * each, a multiple of 4), in point order -- the StreamCompact of the
 * reference (DispatcherGenerateTopology.h:97-110) done with the prepare
 * step's ranks. */
int dax_hip_topology_compact_point_field(const void* d_in, size_t value_bytes, int64_t n_cells,
                                         int64_t n_points, const int32_t* d_point_mask,
                                         const void* d_workspace, size_t workspace_bytes,
                                         void* d_out, void* stream);

/* VTK legacy ASCII files of an output grid held in host memory: the format
 * of the reference benchmarks' PrintContentsToStream
 * (Benchmarks/MarchingCubes/Pipeline.h:51-98: DATASET POLYDATA, POLYGONS;
 * Benchmarks/Threshold/Pipeline.h:99-150: DATASET UNSTRUCTURED_GRID, CELLS,
 * CELL_TYPES with vtk_cell_type for every cell, e.g. 12 = hexahedron,
 * 10 = tetrahedron), numbers as an ostream writes them by default.
 * h_points [n_points][3] f32, h_conn [n_cells][verts_per_cell] int32 (each
 * id in [0, n_points)).  Host code; formats on all host threads. */
int dax_hip_write_vtk_polydata(const char* path, const float* h_points, int64_t n_points,
                               const int32_t* h_conn, int64_t n_cells, int verts_per_cell);
int dax_hip_write_vtk_unstructured(const char* path, const float* h_points, int64_t n_points,
                                   const int32_t* h_conn, int64_t n_cells, int verts_per_cell,
                                   int vtk_cell_type);

/* Synchronous: reads {T, U} written by a forward (or classify) on `stream`,
 * and returns DAX_HIP_CAPACITY if the generate step was skipped for lack of
 * room, DAX_HIP_EXEC_ERROR if T or U left the int32 dax::Id range. */
int dax_hip_mc_read_counts(const int64_t* d_counts, int64_t cap_points, int64_t cap_tris,
                           int64_t* h_tris, int64_t* h_points, void* stream);

/* The three phases of dax_hip_mc_forward as separate calls on one
 * workspace, in this order:
 *   dax_hip_mc_bits      classify: field -> inside-bits (one kernel; also
 *                        resets the scan state of the workspace)
 *   dax_hip_mc_scan      per-word vertex / triangle counts, then the scan
 *                        of the 256-word count tiles' sums (two kernels)
 *   dax_hip_mc_generate  outputs, as dax_hip_mc_forward (d_counts written),
 *                        and every row word's vertex / triangle base in the
 *                        workspace (one kernel)
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
 * params5 = {classify_unroll, classify_vec, scan_sub, gen_tile, gen_xcd}:
 *   classify_vec 1 (one float per lane, unroll 8/16/32 row words per wave)
 *   or 4 (float4 per lane, unroll 2/4/8 chunks of 256 points; rows of a
 *   multiple of 64 points); scan_sub 1/2/4/8 (waves per block of the count
 *   kernel, two count tiles of 128 row words per wave); gen_tile
 *   64/128/256/512/1024 (row words per generate block); gen_xcd 1 gives each
 *   XCD a contiguous range of generate tiles, 0 keeps the plain block order.
 * A shape of (0,0,0) sets or gets the default for shapes not tuned. */
int dax_hip_mc_set_tuning(int64_t nx, int64_t ny, int64_t nz, const int* h_params5);
int dax_hip_mc_get_tuning(int64_t nx, int64_t ny, int64_t nz, int* h_params5);
/* Times the classify variants with HIP events (reps launches each), then
 * every (scan_sub, gen_tile, gen_xcd) combination as count/scan + generate
 * (reps repetitions, each after a classify; the count block size and the
 * tile size together decide the generate tiles' dispatch order), keeps the
 * fastest for this shape and returns it (h_best5) with the milliseconds of
 * classify, count/scan and generate (h_best_ms3, may be NULL).
 * Synchronous.  The outputs and workspace are those of dax_hip_mc_forward. */
int dax_hip_mc_autotune(const float* d_field, int64_t nx, int64_t ny, int64_t nz, float iso,
                        unsigned flags, void* d_workspace, size_t workspace_bytes,
                        float* d_vertices, float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                        int64_t cap_points, int64_t cap_tris, int64_t* d_counts, int reps,
                        int* h_best5, float* h_best_ms3, void* stream);
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

/* ---- batched marching cubes (no reference counterpart) ------------------
 * Many independent volumes: volume v has point dims h_dims3[3v..3v+2],
 * its field at d_fields + h_field_offsets[v] (float32, flat i + nx*(j+ny*k)),
 * isovalue h_isos[v], origin/spacing h_origins3/h_spacings3 (NULL = 0 / 1).
 * Outputs are concatenated in volume order; connectivity is local to each
 * volume; d_vol_offsets [n][2] receives {first triangle, first point} per
 * volume and d_counts {T, U} the totals.  Per volume the results equal a
 * single-volume dax_hip_mc_forward.
 *   dax_hip_mc_batch_workspace_bytes  size for this set of volume dims
 *   dax_hip_mc_batch_prepare          uploads the volume and tile tables
 *                                     (synchronous); copies a header of
 *                                     dax_hip_mc_batch_header_bytes() bytes
 *                                     to h_header_out for the calls below
 *   dax_hip_mc_batch_forward          four persistent launches (classify,
 *                                     count/scan, generate, offsets)
 *   dax_hip_mc_batch_graph_create     the same, captured into a hipGraphExec
 *                                     on a non-null stream
 *   dax_hip_graph_launch / _destroy   replay / free a captured graph */
size_t dax_hip_mc_batch_workspace_bytes(int64_t n_volumes, const int64_t* h_dims3);
size_t dax_hip_mc_batch_header_bytes(void);
int dax_hip_mc_batch_prepare(int64_t n_volumes, const int64_t* h_dims3,
                             const int64_t* h_field_offsets, const float* h_isos,
                             const float* h_origins3, const float* h_spacings3, void* d_workspace,
                             size_t workspace_bytes, void* h_header_out, void* stream);
int dax_hip_mc_batch_forward(const float* d_fields, const void* h_header, void* d_workspace,
                             size_t workspace_bytes, unsigned flags, float* d_vertices,
                             float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                             int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                             int64_t* d_vol_offsets, void* stream);
int dax_hip_mc_batch_graph_create(const float* d_fields, const void* h_header, void* d_workspace,
                                  size_t workspace_bytes, unsigned flags, float* d_vertices,
                                  float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                                  int64_t cap_points, int64_t cap_tris, int64_t* d_counts,
                                  int64_t* d_vol_offsets, void* stream, void** h_graph_exec);
/* Launch shapes of the batch forward, held in the header:
 * {classify unroll 8|16|32, classify blocks per CU 2|4|8, scan threads per
 * 1024-word tile 256|512|1024, generate blocks per CU 2|4|8, generate
 * XCD-contiguous tile order 0|1}.  prepare starts from the shape cached for
 * the batch's class (volume count, ceil(log2(row words)), widest row) or the
 * default {16, 4, 256, 4, 1}.  Every shape gives the same outputs.
 *   dax_hip_mc_batch_set_tuning / _get_tuning  write / read the header's
 *   dax_hip_mc_batch_autotune  times each phase's shapes on this batch (reps
 *       launches each, synchronous), keeps the fastest in the header and the
 *       class cache; h_best5 the choice, h_best_ms3 its classify, count/scan
 *       and generate times.  Outputs as for dax_hip_mc_batch_forward.
 * No reference counterpart (the reference runs one volume per call). */
int dax_hip_mc_batch_set_tuning(void* h_header, const int* h_params5);
int dax_hip_mc_batch_get_tuning(const void* h_header, int* h_params5);
int dax_hip_mc_batch_autotune(const float* d_fields, void* h_header, void* d_workspace,
                              size_t workspace_bytes, unsigned flags, float* d_vertices,
                              float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                              int64_t cap_points, int64_t cap_tris, int64_t* d_counts, int reps,
                              int* h_best5, float* h_best_ms3, void* stream);
/* Batched adjoint: dL/dfield of every volume (written at the fields'
 * offsets) from dL/dvertex and dL/dnormal [U][3] of the concatenated outputs
 * of the last dax_hip_mc_batch_forward on this workspace (run with the edge
 * list).  d_scratch holds dax_hip_mc_backward_scratch_bytes(n_points).
 * No reference counterpart. */
int dax_hip_mc_batch_backward(const float* d_fields, const void* h_header,
                              const void* d_workspace, size_t workspace_bytes,
                              const uint32_t* d_edges, const int64_t* d_counts, int64_t n_points,
                              const float* d_dl_dvertices, const float* d_dl_dnormals,
                              void* d_scratch, size_t scratch_bytes, float* d_dl_dfields,
                              void* stream);

/* Baseline of the batched path: the same volumes as a loop of
 * dax_hip_mc_forward (and, when d_dl_dfields is non-null,
 * dax_hip_mc_backward) calls, one per volume, into the concatenated outputs
 * at h_vol_offsets [n][2] (as the batched forward reports them).
 * d_workspace holds the largest volume's dax_hip_mc_workspace_bytes;
 * d_counts_scratch holds 2 int64.  No reference counterpart. */
int dax_hip_mc_volume_loop(int64_t n_volumes, const int64_t* h_dims3,
                           const int64_t* h_field_offsets, const float* h_isos,
                           const float* h_origins3, const float* h_spacings3,
                           const float* d_fields, unsigned flags, void* d_workspace,
                           size_t workspace_bytes, const int64_t* h_vol_offsets,
                           int64_t cap_points, int64_t cap_tris, float* d_vertices,
                           float* d_normals, int32_t* d_conn, uint32_t* d_edges,
                           int64_t* d_counts_scratch, const float* d_dl_dvertices,
                           const float* d_dl_dnormals, void* d_scratch, size_t scratch_bytes,
                           float* d_dl_dfields, void* stream);

/* Capture every launch this library makes on `stream` between the two calls
 * into one executable graph (e.g. a batch forward and its backward). */
int dax_hip_graph_begin_capture(void* stream);
int dax_hip_graph_end_capture(void* stream, void** h_graph_exec);
int dax_hip_graph_launch(void* graph_exec, void* stream);
int dax_hip_graph_destroy(void* graph_exec);

/* ---- adjoint (no reference counterpart) ---------------------------------
 * dL/dfield of L(vertices, normals) given dL/dvertices [U][3] and
 * dL/dnormals [U][3], with the topology of the forward held fixed.  Uses the
 * forward's workspace (inside-bits and vertex bases must still hold the
 * forward of this field and iso; the backward writes only the dispatch order
 * of its gather tiles there) and its compact edge list d_edges; the
 * interpolants (t, endpoint gradients) are recomputed from the field.
 * U = min(d_counts[1], n_points) if d_counts is given (device {T, U} of the
 * forward), else n_points.  d_scratch needs
 * dax_hip_mc_backward_scratch_bytes(n_points) bytes (a 32-byte record per
 * vertex and a list of the tiles dense enough to be redone in spans).
 * Deterministic: every sum has a fixed order (an atomic counter only fills
 * the tile list, whose order does not change any value), so the result is
 * bitwise reproducible.  Rows of at most 15872 points (nx).  d_dl_dfield [P]
 * f32 is fully written. */
size_t dax_hip_mc_backward_scratch_bytes(int64_t n_points);
int dax_hip_mc_backward(const float* d_field, int64_t nx, int64_t ny, int64_t nz,
                        const float* h_origin3, const float* h_spacing3, float iso,
                        void* d_workspace, size_t workspace_bytes, const uint32_t* d_edges,
                        const int64_t* d_counts, int64_t n_points, const float* d_dl_dvertices,
                        const float* d_dl_dnormals, void* d_scratch, size_t scratch_bytes,
                        float* d_dl_dfield, void* stream);

/* No reference counterpart.  Gather tiles too dense for their block's
 * endpoint records are redone in spans: by their own block, one span after
 * another, when the volume has at least `tiles_per_cu` gather tiles per CU
 * (many rounds of tiles hide such a block), else in a separate span pass
 * over many blocks.  Default 16; 0 keeps the spans in their tile's block at
 * any size.  Process-wide; the result does not depend on it (bit for bit). */
int dax_hip_mc_backward_set_span_threshold(int64_t tiles_per_cu);

/* The same for a Vector3 point field ([n][3] f32, e.g. coordinates). */
int dax_hip_mc_interpolate_vec3(const float* d_in, const int32_t* d_first,
                                const int32_t* d_second, const float* d_weight, int64_t n,
                                float* d_out, void* stream);

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
