/*
 * C ABI of the MI355X-native LUE flow-routing path (liblue_amd.so).
 *
 * Each entry point replaces one explicit instantiation of LUE's routing operations for
 * DefaultValuePolicies (ldd uint8, material float | double), which LUE exports from
 * liblue_framework_routing_operation (reference: source/framework/algorithm/CMakeLists.txt:1980-2025,
 * source/framework/algorithm/source/routing_operation/ *.cpp.in templates). The C++ front end
 * (lue_amd/include/lue/framework/...) maps lue::value_policies::* and lue::* calls on
 * PartitionedArray onto these functions.
 *
 * Conventions
 *   - All raster pointers are DEVICE pointers to whole arrays, row-major, rows x cols cells
 *     (partitions are contiguous sub-blocks in LUE; the partition grid changes no result except the cell
 *     classes of accu_info, which takes the partition shape as arguments).
 *   - A raster argument may be NULL; the argument is then the Scalar that follows it, broadcast to
 *     every cell (LUE's Scalar<E> overloads, accumulating_router.hpp:727-786).
 *   - `stream` is a hipStream_t (NULL: the default stream). Every call enqueues its work on `stream` and
 *     returns without waiting for it: the outputs are complete when the stream reaches that point
 *     (synchronise the stream, or make other streams wait for an event recorded after the call, before
 *     reading them). The per-call routing entry points (accu, accu_threshold, accu_fraction,
 *     accu_capacity, accu_trigger, kinematic_wave) first compute a fingerprint of the ldd bytes on
 *     `stream` and wait for it, which also waits for the work queued on `stream` before the call; when
 *     they record a new schedule for the ldd they wait for the recording as well (so an invalid network
 *     is reported by the call that meets it). Callers that know when their ldd changes avoid that wait
 *     with a flow network (lue_amd_network_*), which never waits.
 *   - No-data: ldd 255 (numeric_limits<uint8>::max()), float NaN. Cells whose inputs are no-data or
 *     outside the operation's domain produce NaN, as do all cells downstream of them.
 *   - Return value: LUE_AMD_SUCCESS or an error status; lue_amd_last_error() gives a message.
 */
#ifndef LUE_AMD_ROUTING_H
#define LUE_AMD_ROUTING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUE_AMD_SUCCESS 0
#define LUE_AMD_ERROR_INVALID_ARGUMENT 1
#define LUE_AMD_ERROR_SHAPE_MISMATCH 2
#define LUE_AMD_ERROR_OUT_OF_MEMORY 3
#define LUE_AMD_ERROR_HIP 4
#define LUE_AMD_ERROR_INVALID_NETWORK 5
#define LUE_AMD_ERROR_NO_DEVICE 6

/* Message for a status code, and the detail message of the last failing call on this thread. */
const char* lue_amd_status_string(int status);
const char* lue_amd_last_error(void);

/* Library version, "major.minor.patch". */
const char* lue_amd_version(void);

/*
 * Timing of the routing sweeps with HIP events on the caller's stream (no reference counterpart; it plays
 * the role of LUE's QA stopwatch instrumentation, source/quality_assurance/cxx). enable != 0 resets and
 * starts collection on this process (enable == 0 resets and stops it); the operations record their events
 * without waiting for them. read waits for the events recorded so far and returns the summed time of the
 * first (all-tile) rounds, of the later rounds, the number of routing calls, of rounds and of cells;
 * read_op the same for one operation (LUE_AMD_OP_*) and element type (f64 != 0: double), with the
 * shortest and longest first round.
 */
#define LUE_AMD_OP_ACCU 0
#define LUE_AMD_OP_ACCU_THRESHOLD 1
#define LUE_AMD_OP_ACCU_FRACTION 2
#define LUE_AMD_OP_KINEMATIC_WAVE 3
#define LUE_AMD_OP_ACCU_CAPACITY 4
#define LUE_AMD_OP_ACCU_TRIGGER 5
int lue_amd_profile_enable(int enable);
int lue_amd_profile_read(double* first_round_ms, double* next_rounds_ms, long long* calls, long long* rounds,
                         long long* cells);
int lue_amd_profile_read_op(int op, int f64, double* first_round_ms, double* next_rounds_ms,
                            double* first_round_min_ms, double* first_round_max_ms, long long* calls);

/*
 * Schedules of the routing sweep after its first round over all tiles (no reference counterpart):
 * LUE_AMD_SCHEDULE_CHAINS (default): the cells the first round leaves pending are solved along heavy-path
 * chains, one launch per chain level; LUE_AMD_SCHEDULE_ROUNDS: further rounds over the tiles that received
 * flow until none is left. Results do not depend on the schedule beyond floating point summation order.
 */
#define LUE_AMD_SCHEDULE_CHAINS 0
#define LUE_AMD_SCHEDULE_ROUNDS 1

/* Schedule of the per-call entry points (default LUE_AMD_SCHEDULE_CHAINS). */
int lue_amd_set_sweep_options(int schedule);

/*
 * Per-call entry points keep the schedules they record (no reference counterpart; LUE recomputes the
 * inflow counts in every call): up to `entries` of them (default 4, 0 disables), found again by device,
 * ldd pointer, shape, tile shape, schedule and a 128-bit fingerprint of the ldd bytes computed in every
 * call, so an ldd changed in place is recorded anew.
 */
int lue_amd_set_network_cache(int entries);

/*
 * Per-call device workspaces come from a cache of device blocks that later calls reuse (a block released
 * on a stream is reused on that stream at once, on another once its last use has completed). This frees
 * the idle blocks of the current device, waiting for their last use. No counterpart in the reference
 * (LUE allocates per partition task).
 */
int lue_amd_trim_scratch(void);

/* Device memory used by the C++ front end's device-resident partitions. The copies wait for `stream`. */
int lue_amd_device_malloc(void** ptr, size_t bytes);
int lue_amd_device_free(void* ptr);
int lue_amd_copy_to_device(void* dst, const void* src, size_t bytes, void* stream);
int lue_amd_copy_to_host(void* dst, const void* src, size_t bytes, void* stream);
int lue_amd_stream_synchronize(void* stream);

/*
 * Completion events (hipEvent_t) of the C++ front end's partitioned arrays, which stand in for the futures
 * of LUE's partitions (source/framework/partitioned_array/include/lue/framework/partitioned_array_decl.hpp:
 * 23-98, partitions are hpx::future-backed components): an operation records one after its launches, later
 * operations make their stream wait for it, host reads wait for it. query sets *complete to 1 when the
 * event has completed, 0 otherwise.
 */
int lue_amd_event_create(void** event);
int lue_amd_event_destroy(void* event);
int lue_amd_event_record(void* event, void* stream);
int lue_amd_stream_wait_event(void* stream, void* event);
int lue_amd_event_synchronize(void* event);
int lue_amd_event_query(void* event, int* complete);

/*
 * Inflow count: number of upstream neighbours per cell, 255 where the ldd is no-data.
 * Replaces lue::inflow_count<std::uint8_t>(policies, flow_direction)
 * (ALG/inflow_count.hpp:29-42, ALG/definition/inflow_count.hpp:665-726, instantiated for uint8 at
 * ALG/definition/inflow_count.hpp:731-735). ALG = source/framework/algorithm/include/lue/framework/algorithm
 */
int lue_amd_inflow_count_u8(const uint8_t* ldd, int64_t rows, int64_t cols, uint8_t* inflow_count, void* stream);

/*
 * accu_info: the class LUE's accumulating router leaves in each cell when the array is split into
 * partitions of partition_rows x partition_cols cells (floor(rows / partition_rows) partitions per column,
 * the last one absorbing the remainder; the same along rows): 11 ridge, 12 intra-partition stream,
 * 13 partition output, 14 partition input, 15 partition input and output, 16 inter-partition stream,
 * 17 sink, 255 where the ldd is no-data. A partition input cell that also has an upstream neighbour in
 * its own partition waiting for another partition gets 14/15; in LUE it can also get 16/13/17, depending on
 * which arrives last. Runs accu (float) of a crossing indicator on the per-call network of the ldd.
 * Replaces lue::value_policies::accu_info(flow_direction) (ALG/value_policies/accu_info.hpp:19-28,
 * ALG/accu_info.hpp:22-78, ALG/definition/accu_info.hpp:10-188; Python binding
 * source/framework/python/source/algorithm/routing_operation/accu_info.cpp:10-17).
 * Asynchronous on `stream`.
 */
int lue_amd_accu_info_u8(const uint8_t* ldd, int64_t rows, int64_t cols, int64_t partition_rows,
                         int64_t partition_cols, uint8_t* cell_class, void* stream);

/*
 * d8_flow_direction: the ldd of an elevation raster (NaN = no-data -> 255). Each cell drains to the
 * neighbour with the largest raw drop (no distance weighting; first in the order N, W, E, S, NW, NE, SW,
 * SE on ties; sink when no drop is positive); outside the array the elevation is numeric_limits<T>::min().
 * Replaces lue::value_policies::d8_flow_direction<uint8_t>(elevation)
 * (ALG/value_policies/d8_flow_direction.hpp:9-57, ALG/definition/d8_flow_direction.hpp:12-173).
 * Asynchronous on `stream`.
 */
int lue_amd_d8_flow_direction_f32(const float* elevation, int64_t rows, int64_t cols, uint8_t* flow_direction,
                                  void* stream);
int lue_amd_d8_flow_direction_f64(const double* elevation, int64_t rows, int64_t cols, uint8_t* flow_direction,
                                  void* stream);

/*
 * downstream_distance: per cell, the distance to its downstream cell: cell_size for N, W, E, S,
 * sqrt(2) * cell_size (rounded in the element type) for NW, NE, SW, SE, 0 for a pit or unknown code. With
 * mark_no_data != 0 (LUE's value policies, as lue.framework.downstream_distance uses them) a no-data ldd
 * (255) gives NaN; with 0 (lue::downstream_distance's default policies) it gives 0. flow_direction must be
 * 4-byte aligned. Replaces lue::downstream_distance (ALG/downstream_distance.hpp:24-39,
 * ALG/definition/downstream_distance.hpp:13-75; Python binding
 * source/framework/python/source/algorithm/routing_operation/downstream_distance.cpp:14-23).
 * Asynchronous on `stream`.
 */
int lue_amd_downstream_distance_f32(const uint8_t* flow_direction, int64_t rows, int64_t cols, float cell_size,
                                   int mark_no_data, float* distance, void* stream);
int lue_amd_downstream_distance_f64(const uint8_t* flow_direction, int64_t rows, int64_t cols, double cell_size,
                                   int mark_no_data, double* distance, void* stream);

/*
 * upstream: per cell, the sum (from 0, in the window's row-major order: NW, N, NE, W, E, SW, S, SE
 * neighbour) of the material of the neighbours that drain into it; downstream: per cell, the material of
 * its downstream neighbour (0 beyond the array's edge; its own for a pit). A no-data ldd (255) or material
 * in the cell, or a no-data material taken from a neighbour, gives no-data: NaN for floating point,
 * the minimum of a signed and the maximum of an unsigned integer type. Integer sums wrap.
 * Replaces lue::value_policies::upstream / downstream (ALG/value_policies/upstream.hpp:16-30,
 * ALG/definition/upstream.hpp:36-190; ALG/value_policies/downstream.hpp:16-30,
 * ALG/definition/downstream.hpp:36-186) for LUE's default material element types
 * (environment/cmake/LueConfiguration.cmake:171-228). Asynchronous on `stream`.
 */
int lue_amd_upstream_u8(const uint8_t* flow_direction, const uint8_t* material, int64_t rows, int64_t cols,
                        uint8_t* out, void* stream);
int lue_amd_upstream_u32(const uint8_t* flow_direction, const uint32_t* material, int64_t rows, int64_t cols,
                         uint32_t* out, void* stream);
int lue_amd_upstream_u64(const uint8_t* flow_direction, const uint64_t* material, int64_t rows, int64_t cols,
                         uint64_t* out, void* stream);
int lue_amd_upstream_i8(const uint8_t* flow_direction, const int8_t* material, int64_t rows, int64_t cols,
                        int8_t* out, void* stream);
int lue_amd_upstream_i32(const uint8_t* flow_direction, const int32_t* material, int64_t rows, int64_t cols,
                         int32_t* out, void* stream);
int lue_amd_upstream_i64(const uint8_t* flow_direction, const int64_t* material, int64_t rows, int64_t cols,
                         int64_t* out, void* stream);
int lue_amd_upstream_f32(const uint8_t* flow_direction, const float* material, int64_t rows, int64_t cols,
                         float* out, void* stream);
int lue_amd_upstream_f64(const uint8_t* flow_direction, const double* material, int64_t rows, int64_t cols,
                         double* out, void* stream);
int lue_amd_downstream_u8(const uint8_t* flow_direction, const uint8_t* material, int64_t rows, int64_t cols,
                          uint8_t* out, void* stream);
int lue_amd_downstream_u32(const uint8_t* flow_direction, const uint32_t* material, int64_t rows, int64_t cols,
                           uint32_t* out, void* stream);
int lue_amd_downstream_u64(const uint8_t* flow_direction, const uint64_t* material, int64_t rows, int64_t cols,
                           uint64_t* out, void* stream);
int lue_amd_downstream_i8(const uint8_t* flow_direction, const int8_t* material, int64_t rows, int64_t cols,
                          int8_t* out, void* stream);
int lue_amd_downstream_i32(const uint8_t* flow_direction, const int32_t* material, int64_t rows, int64_t cols,
                           int32_t* out, void* stream);
int lue_amd_downstream_i64(const uint8_t* flow_direction, const int64_t* material, int64_t rows, int64_t cols,
                           int64_t* out, void* stream);
int lue_amd_downstream_f32(const uint8_t* flow_direction, const float* material, int64_t rows, int64_t cols,
                           float* out, void* stream);
int lue_amd_downstream_f64(const uint8_t* flow_direction, const double* material, int64_t rows, int64_t cols,
                           double* out, void* stream);

/*
 * accu: outflow = sum of upstream outflow + inflow.
 * Replaces lue::value_policies::accu(flow_direction, inflow | Scalar inflow)
 * (ALG/value_policies/accu.hpp:33-55, ALG/accu.hpp:12-28, ALG/definition/accu.hpp:226-264).
 */
int lue_amd_accu_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow, float inflow_scalar,
                     float* outflow, void* stream);
int lue_amd_accu_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                     double inflow_scalar, double* outflow, void* stream);

/*
 * accu_threshold: material above the threshold flows on, the rest stays (remainder).
 * Replaces lue::value_policies::accu_threshold(flow_direction, inflow, threshold), all four
 * raster/Scalar combinations (ALG/value_policies/accu_threshold.hpp:34-87, ALG/accu_threshold.hpp:9-50,
 * ALG/definition/accu_threshold.hpp:29-395).
 */
int lue_amd_accu_threshold_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                               float inflow_scalar, const float* threshold, float threshold_scalar, float* outflow,
                               float* remainder, void* stream);
int lue_amd_accu_threshold_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                               double inflow_scalar, const double* threshold, double threshold_scalar,
                               double* outflow, double* remainder, void* stream);

/*
 * accu_fraction: a fraction of the material flows on, the rest stays (remainder).
 * Replaces lue::value_policies::accu_fraction(flow_direction, inflow, fraction), all four
 * raster/Scalar combinations (ALG/value_policies/accu_fraction.hpp, ALG/accu_fraction.hpp:9-50,
 * ALG/definition/accu_fraction.hpp:29-387).
 */
int lue_amd_accu_fraction_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                              float inflow_scalar, const float* fraction, float fraction_scalar, float* outflow,
                              float* remainder, void* stream);
int lue_amd_accu_fraction_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                              double inflow_scalar, const double* fraction, double fraction_scalar,
                              double* outflow, double* remainder, void* stream);

/*
 * accu_capacity: at most the capacity flows on, the rest stays (remainder).
 * Replaces lue::value_policies::accu_capacity(flow_direction, inflow, capacity), all four
 * raster/Scalar combinations (ALG/value_policies/accu_capacity.hpp:34-88, ALG/accu_capacity.hpp:9-50,
 * ALG/definition/accu_capacity.hpp:124-166).
 */
int lue_amd_accu_capacity_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                              float inflow_scalar, const float* capacity, float capacity_scalar, float* outflow,
                              float* remainder, void* stream);
int lue_amd_accu_capacity_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                              double inflow_scalar, const double* capacity, double capacity_scalar,
                              double* outflow, double* remainder, void* stream);

/*
 * accu_trigger: material below the trigger stays (remainder), from the trigger on all of it flows on.
 * Replaces lue::value_policies::accu_trigger(flow_direction, inflow, trigger), all four
 * raster/Scalar combinations (ALG/value_policies/accu_trigger.hpp:34-88, ALG/accu_trigger.hpp:9-50,
 * ALG/definition/accu_trigger.hpp:124-164).
 */
int lue_amd_accu_trigger_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                             float inflow_scalar, const float* trigger, float trigger_scalar, float* outflow,
                             float* remainder, void* stream);
int lue_amd_accu_trigger_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                             double inflow_scalar, const double* trigger, double trigger_scalar,
                             double* outflow, double* remainder, void* stream);

/*
 * kinematic_wave: new discharge per cell from the upstream new discharge, the current discharge and
 * the lateral inflow (non-linear kinematic wave, Newton solve per cell).
 * Replaces both lue::value_policies::kinematic_wave overloads (ALG/value_policies/kinematic_wave.hpp:43-91,
 * ALG/kinematic_wave.hpp:9-32, ALG/definition/kinematic_wave.hpp:606-684): current_outflow and inflow
 * are rasters; alpha, beta, channel_length are rasters or Scalars; time_step_duration is a Scalar.
 * f64: Newton in FP64 (LUE's algorithm). f32: Newton in FP32 plus one FP64 correction per cell, upstream
 * sums carried in compensated form.
 */
int lue_amd_kinematic_wave_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* current_outflow,
                               const float* inflow, const float* alpha, float alpha_scalar, const float* beta,
                               float beta_scalar, float time_step_duration, const float* channel_length,
                               float channel_length_scalar, float* outflow, void* stream);
int lue_amd_kinematic_wave_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* current_outflow,
                               const double* inflow, const double* alpha, double alpha_scalar, const double* beta,
                               double beta_scalar, double time_step_duration, const double* channel_length,
                               double channel_length_scalar, double* outflow, void* stream);

/*
 * Kinematic-wave Newton solve of single cells, elementwise over n inputs (device pointers).
 * Replaces lue::detail::iterate_to_new_discharge<T> (ALG/definition/kinematic_wave.hpp:145-254),
 * including extraction for negative lateral inflow; used by the unit tests of the solve.
 */
int lue_amd_iterate_to_new_discharge_f32(const float* upstream_discharge, const float* current_discharge,
                                         const float* lateral_inflow, const float* alpha, const float* beta,
                                         const float* time_step_duration, const float* channel_length,
                                         float* new_discharge, int64_t n, void* stream);
int lue_amd_iterate_to_new_discharge_f64(const double* upstream_discharge, const double* current_discharge,
                                         const double* lateral_inflow, const double* alpha, const double* beta,
                                         const double* time_step_duration, const double* channel_length,
                                         double* new_discharge, int64_t n, void* stream);

/*
 * Seeded synthetic inputs of the benchmark workloads (no reference counterpart: LUE's own tests and
 * benchmarks build inputs in C++). ldd from the DEM e(r,c) = (rows - r) + 50 |sin(pi c / 256)| + 0.9 u,
 * sqrt(2)-weighted steepest descent, bottom row outlets; uniform fields lo + (hi - lo) u. The ldd
 * generators return at once (stream order).
 */
int lue_amd_synthetic_ldd(uint64_t seed, int64_t rows, int64_t cols, uint8_t* ldd, void* stream);
/* the DEM e(r, c) itself, rounded to the output type (stream order) */
int lue_amd_synthetic_dem_f32(uint64_t seed, int64_t rows, int64_t cols, float* out, void* stream);
int lue_amd_synthetic_dem_f64(uint64_t seed, int64_t rows, int64_t cols, double* out, void* stream);
int lue_amd_synthetic_uniform_f32(uint64_t seed, uint64_t stream_id, double lo, double hi, int64_t n, float* out,
                                  void* stream);
int lue_amd_synthetic_uniform_f64(uint64_t seed, uint64_t stream_id, double lo, double hi, int64_t n, double* out,
                                  void* stream);

/*
 * Flow networks (no reference counterpart; LUE recomputes the inflow counts of an ldd in every call).
 * A network is a batch of n catchments (independent ldd rasters, device pointers, kept borrowed: they
 * must outlive the network and must not change) together with the schedule of the routing sweep, which
 * depends on the ldd only (the chain layout, or the tiles of each round). Operations on a network
 * run that schedule as a fixed sequence of launches on `stream`, for all catchments at once, and return
 * without waiting for the results (synchronise the stream before reading them).
 * tile_shape: LUE_AMD_TILE_* (width x height cells per tile; operations that do not fit the shape in LDS
 * return LUE_AMD_ERROR_INVALID_ARGUMENT). The per-call entry points use 32x32 for the kinematic wave and
 * 64x32 for the accu family. schedule: LUE_AMD_SCHEDULE_*. Creating a network validates it (cycles:
 * LUE_AMD_ERROR_INVALID_NETWORK).
 * The arguments of the operations are host arrays with one device pointer per catchment; an argument
 * array may be NULL where the operation accepts a Scalar, which then applies to every catchment.
 */
#define LUE_AMD_TILE_32X32 0
#define LUE_AMD_TILE_64X32 1
#define LUE_AMD_TILE_64X64 2
#define LUE_AMD_TILE_32X64 3

typedef struct lue_amd_network_s* lue_amd_network;

int lue_amd_network_create(int n, const uint8_t* const* ldd, const int64_t* rows, const int64_t* cols, int tile_shape,
                           int schedule, void* stream, lue_amd_network* network);
int lue_amd_network_destroy(lue_amd_network network);
/* info[0..16]: catchments, cells, tiles, recorded rounds, cells left pending by the first round, chains,
 * chain levels, no-data cells, tile shape, schedule, longest chain, side references, 64-position blocks
 * the first round runs on the in-tile layouts, and of them the blocks on the tiles' critical paths, the
 * wide threshold (lue_amd_network_set_wide_threshold), the tile queue (lue_amd_network_set_tile_queue), the
 * lane limit (lue_amd_network_set_lane_limit) */
int lue_amd_network_info(lue_amd_network network, int64_t* info, int n);

/*
 * The wave schedule of the chain levels (no reference counterpart; a tuning knob of the batching addition):
 * a level whose longest work item has at least `positions` chain positions runs one work item per workgroup
 * of 16 wavefronts, 1024 positions at a time (accu, threshold, capacity, fraction, kinematic wave; the same
 * results as one wavefront per work item); 0 runs every level one wavefront per work item. Default 2048.
 * Applies to operations enqueued from then on (plans keep the schedule they captured); not while another
 * thread enqueues operations on the network.
 */
int lue_amd_network_set_wide_threshold(lue_amd_network network, int64_t positions);

/*
 * The launch of the chain schedule's first round (no reference counterpart; a tuning knob of the batching
 * addition): on = 0 launches one workgroup per tile of the batch (default), on != 0 a persistent launch of
 * as many workgroups as the device holds at once, fed by a tile work queue (one atomic counter). The results
 * are the same. Applies to operations enqueued from then on (plans keep the launch they captured); not while
 * another thread enqueues operations on the network.
 */
int lue_amd_network_set_tile_queue(lue_amd_network network, int on);

/*
 * The lane schedule of the chain levels (no reference counterpart; a tuning knob of the batching addition):
 * work items of at most `positions` chain positions are walked one per lane, cell after cell (each cell
 * solved once from its known upstream outflow, as in LUE's walk), the longer ones one per wavefront, 64
 * positions at a time (or, in a level on the wide path, one per workgroup); 0 (the default) walks none. Whether
 * an item is walked depends on its length only, not on the other catchments of a batch. The
 * results agree with the other schedule to the operations' tolerances (sums in another order; the kinematic
 * wave's roots to within its FP64 correction). Applies to operations enqueued from then on (plans keep the
 * schedule they captured); not while another thread enqueues operations on the network.
 */
int lue_amd_network_set_lane_limit(lue_amd_network network, int64_t positions);

/*
 * One chain level of a network (chain schedule; no reference counterpart): info[0] chain positions,
 * [1] work items, [2] positions of the longest work item, [3] side references gathered before the level.
 * level is 0 .. levels - 1 (lue_amd_network_info: info[6]).
 */
int lue_amd_network_level_info(lue_amd_network network, int level, int64_t* info, int n);

int lue_amd_network_accu_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                             float* const* outflow, void* stream);
int lue_amd_network_accu_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                             double* const* outflow, void* stream);
int lue_amd_network_accu_threshold_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                       const float* const* threshold, float threshold_scalar, float* const* outflow,
                                       float* const* remainder, void* stream);
int lue_amd_network_accu_threshold_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                       const double* const* threshold, double threshold_scalar,
                                       double* const* outflow, double* const* remainder, void* stream);
int lue_amd_network_accu_fraction_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                      const float* const* fraction, float fraction_scalar, float* const* outflow,
                                      float* const* remainder, void* stream);
int lue_amd_network_accu_fraction_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                      const double* const* fraction, double fraction_scalar, double* const* outflow,
                                      double* const* remainder, void* stream);
int lue_amd_network_accu_capacity_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                      const float* const* capacity, float capacity_scalar, float* const* outflow,
                                      float* const* remainder, void* stream);
int lue_amd_network_accu_capacity_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                      const double* const* capacity, double capacity_scalar, double* const* outflow,
                                      double* const* remainder, void* stream);
int lue_amd_network_accu_trigger_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                     const float* const* trigger, float trigger_scalar, float* const* outflow,
                                     float* const* remainder, void* stream);
int lue_amd_network_accu_trigger_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                     const double* const* trigger, double trigger_scalar, double* const* outflow,
                                     double* const* remainder, void* stream);
int lue_amd_network_kinematic_wave_f32(lue_amd_network network, const float* const* current_outflow,
                                       const float* const* inflow, const float* const* alpha, float alpha_scalar,
                                       const float* const* beta, float beta_scalar, float time_step_duration,
                                       const float* const* channel_length, float channel_length_scalar,
                                       float* const* outflow, void* stream);
int lue_amd_network_kinematic_wave_f64(lue_amd_network network, const double* const* current_outflow,
                                       const double* const* inflow, const double* const* alpha, double alpha_scalar,
                                       const double* const* beta, double beta_scalar, double time_step_duration,
                                       const double* const* channel_length, double channel_length_scalar,
                                       double* const* outflow, void* stream);

/*
 * Plans: one operation on a network with fixed arguments, captured into a HIP graph (no reference
 * counterpart; LUE's timestep loops re-issue the operation's tasks, source/framework/model/include/lue/
 * framework/model/simulate.hpp:60-113). Creating a plan allocates the operation's per-call buffers, writes
 * the per-catchment argument table once and captures the launch sequence (waits for the device); a launch
 * replays the whole sequence as one graph launch on `stream` (asynchronous), reading and writing the rasters
 * given at creation (their contents may change between launches, the pointers may not). The network and the
 * rasters must outlive the plan; launches of one plan must not overlap (they share its buffers). The plan
 * functions take the arguments of the corresponding lue_amd_network_* functions, without the stream.
 * info[0..1]: kernel launches per replay (graph nodes), bytes of per-call buffers.
 */
typedef struct lue_amd_plan_s* lue_amd_plan;

int lue_amd_plan_launch(lue_amd_plan plan, void* stream);
int lue_amd_plan_destroy(lue_amd_plan plan);
int lue_amd_plan_info(lue_amd_plan plan, int64_t* info, int n);

int lue_amd_plan_accu_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                          float* const* outflow, lue_amd_plan* plan);
int lue_amd_plan_accu_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                          double* const* outflow, lue_amd_plan* plan);
int lue_amd_plan_accu_threshold_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                    const float* const* threshold, float threshold_scalar, float* const* outflow,
                                    float* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_accu_threshold_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                    const double* const* threshold, double threshold_scalar, double* const* outflow,
                                    double* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_accu_fraction_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                   const float* const* fraction, float fraction_scalar, float* const* outflow,
                                   float* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_accu_fraction_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                   const double* const* fraction, double fraction_scalar, double* const* outflow,
                                   double* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_accu_capacity_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                   const float* const* capacity, float capacity_scalar, float* const* outflow,
                                   float* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_accu_capacity_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                   const double* const* capacity, double capacity_scalar, double* const* outflow,
                                   double* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_accu_trigger_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                  const float* const* trigger, float trigger_scalar, float* const* outflow,
                                  float* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_accu_trigger_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                  const double* const* trigger, double trigger_scalar, double* const* outflow,
                                  double* const* remainder, lue_amd_plan* plan);
int lue_amd_plan_kinematic_wave_f32(lue_amd_network network, const float* const* current_outflow,
                                    const float* const* inflow, const float* const* alpha, float alpha_scalar,
                                    const float* const* beta, float beta_scalar, float time_step_duration,
                                    const float* const* channel_length, float channel_length_scalar,
                                    float* const* outflow, lue_amd_plan* plan);
int lue_amd_plan_kinematic_wave_f64(lue_amd_network network, const double* const* current_outflow,
                                    const double* const* inflow, const double* const* alpha, double alpha_scalar,
                                    const double* const* beta, double beta_scalar, double time_step_duration,
                                    const double* const* channel_length, double channel_length_scalar,
                                    double* const* outflow, lue_amd_plan* plan);

/*
 * lue_amd_network_kinematic_wave_f32 for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation). No counterpart in the reference. It
 * allocates nothing: its per-call buffers (the replay's staged chain inputs, side sums, chain results and
 * side slots) live in one fixed 64 MiB device array per GPU, held until the process ends, which fits
 * networks of up to about 5.5 million cells on the synthetic DEM (about 11.6 bytes per cell; a
 * larger network returns LUE_AMD_ERROR_INVALID_ARGUMENT with the bytes it needs) and of at most 32
 * catchments. The profiler (lue_amd_profile_*) does not time these calls. Every call on one GPU shares that
 * array: calls that overlap on one GPU (on different streams, or from different threads) race; calls on one
 * stream are ordered. Arguments and results as lue_amd_network_kinematic_wave_f32; asynchronous on `stream`.
 */
int lue_amd_network_kinematic_wave_f32_noalloc(lue_amd_network network, const float* const* current_outflow,
                                               const float* const* inflow, const float* const* alpha,
                                               float alpha_scalar, const float* const* beta, float beta_scalar,
                                               float time_step_duration, const float* const* channel_length,
                                               float channel_length_scalar, float* const* outflow, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LUE_AMD_ROUTING_H */
