/* vortrace_hip.h -- C-ABI of the MI355X (gfx950) ray-projection library.
 *
 * This is the drop-in boundary that replaces the reference's pybind11 module
 * `Cvortrace` (src/bindings.cpp:71-326).  Every entry point takes plain
 * pointers and sizes; arrays marked d_* live in device memory on the device
 * that was current when the cloud was built, all others in host memory.
 * Calls are ordered on the HIP stream passed as `stream` (a hipStream_t, or
 * NULL for the default stream) and return after the work has completed, so
 * that errors raised inside kernels can be reported as status codes.
 *
 * Status codes: VT_OK, VT_EINVAL (the reference throws std::invalid_argument,
 * which its binding maps to ValueError), VT_ERUNTIME (any other std::exception,
 * RuntimeError).  vt_last_error() returns the message of the last failure on
 * the calling thread, with the reference's wording.
 *
 * A cloud owns per-cloud device work memory (the kernels' per-lane stacks,
 * the deferred-ray list, the adjoint's and field update's buffers, kept
 * between calls): calls on one cloud must not overlap in time; different
 * clouds may be used from different threads at once.
 *
 * Cell ids crossing the boundary are indices into the filtered particle list,
 * as in the reference (PointCloud::orig_ids maps them to input rows).
 */
#ifndef VORTRACE_HIP_H
#define VORTRACE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VT_OK = 0, VT_EINVAL = 1, VT_ERUNTIME = 2 };

/* ReductionMode, include/mytypes.hpp:14 (values are part of the ABI). */
enum { VT_SUM = 0, VT_MAX = 1, VT_MIN = 2, VT_VOLUME = 3 };

typedef struct vt_cloud vt_cloud;

/* Message of the last failed call on this thread (never NULL). */
const char *vt_last_error(void);

/* vortrace::warn (include/mytypes.hpp:19-25) routed to a callback; the
 * default prints "vortrace warning: <msg>" to stderr.  NULL restores it. */
typedef void (*vt_warning_fn)(const char *msg, void *user);
void vt_set_warning_callback(vt_warning_fn fn, void *user);

/* set_verbose / get_verbose (src/bindings.cpp:91-95). */
void vt_set_verbose(int verbose);
int vt_get_verbose(void);

/* PointCloud::loadPoints (src/pointcloud.cpp:12-191; binding
 * src/bindings.cpp:29-69,112-116).  pos is npart x 3, fields npart x nfields,
 * both row-major FP64; subbox = {xmin,xmax,ymin,ymax,zmin,zmax}; vol may be
 * NULL (nvol = 0).  Filters to the padded box (non-periodic) or validates
 * the periodic domain, and computes the RGBA validity flag. */
int vt_cloud_create(const double *pos, size_t npart, const double *fields, size_t nfields,
                    const double subbox[6], const double *vol, size_t nvol, int periodic,
                    vt_cloud **out);

/* PointCloud::buildTree (src/pointcloud.cpp:193-212): the nanoflann 1.3.2
 * L=10 static kd-tree, with identical split decisions, flattened into the
 * device layout and uploaded to the current HIP device. */
int vt_cloud_build_tree(vt_cloud *cloud);

/* The two halves of vt_cloud_build_tree: the host-side build (touches no
 * device) and the upload of a built tree to the current HIP device. */
int vt_cloud_build_host_tree(vt_cloud *cloud);
int vt_cloud_upload(vt_cloud *cloud);

/* Copies the host tree out in pre-order (no reference counterpart; used to
 * check the build against nanoflann's index dump, include/nanoflann.hpp:
 * 1023-1071).  vind has npart entries; the node arrays have
 * vt_tree_info.nodes entries: feat = -1 for a leaf with a/b = first/end
 * position in vind, else the cut dimension with a/b = child1/child2 node
 * indices and divlow/divhigh; bbox6 = root box {xlo,xhi,ylo,yhi,zlo,zhi}. */
int vt_cloud_tree_export(const vt_cloud *cloud, int64_t *vind, int32_t *feat, int64_t *a,
                         int64_t *b, double *divlow, double *divhigh, double *bbox6);

void vt_cloud_destroy(vt_cloud *cloud);

/* Getters (src/bindings.cpp:134-154). */
size_t vt_cloud_npart(const vt_cloud *cloud);
size_t vt_cloud_nfields(const vt_cloud *cloud);
int vt_cloud_periodic(const vt_cloud *cloud);
int vt_cloud_valid_rgba(const vt_cloud *cloud);
int vt_cloud_tree_built(const vt_cloud *cloud);
double vt_cloud_pad(const vt_cloud *cloud);
void vt_cloud_subbox(const vt_cloud *cloud, double out6[6]);
/* out has vt_cloud_npart() entries. */
void vt_cloud_orig_ids(const vt_cloud *cloud, int64_t *out);

/* Tree statistics: number of nodes, leaves and the maximum depth (root = 0). */
typedef struct {
  int64_t nodes, leaves, max_depth;
  double build_seconds;
} vt_tree_info;
int vt_cloud_tree_info(const vt_cloud *cloud, vt_tree_info *out);

/* PointCloud::queryTreeK (src/pointcloud.cpp:308-368, binding 155-174) for nq
 * device-resident query points: ids/r2 are nq x k, sorted by r^2 with
 * nanoflann's tie order.  1 <= k <= 128.  Test hook and building block. */
int vt_query_knn(const vt_cloud *cloud, const double *d_q, size_t nq, int k, int64_t *d_ids,
                 double *d_r2, void *stream);

/* Projection::makeProjection (src/projection.cpp:26-73): one ray per row of
 * d_start/d_end (nrays x 3), result d_out nrays x n_out with n_out = 3 for
 * VT_VOLUME and nfields otherwise (src/reduction.cpp:98-101). */
int vt_project(const vt_cloud *cloud, const double *d_start, const double *d_end, size_t nrays,
               int mode, double *d_out, void *stream);
/* vt_project that allocates nothing: every per-call device buffer (the
 * deferred-ray list, the kNN stacks below their LDS windows, the walk's
 * pending split points, the K_EXPAND buffers) comes out of one static
 * 128 MiB device array that the library holds on each GPU it runs on for the
 * life of the process, and the call's error record and statistics are
 * cleared by a kernel, not a memset.  The caller passes the inputs and the
 * output as for vt_project; the persistent grid is as wide as the array
 * holds buffers for.  At most 4,194,304 rays per call (VT_EINVAL beyond).
 * Calls that overlap in time on one GPU share the array and race.  No
 * reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). */
int vt_project_noalloc(const vt_cloud *cloud, const double *d_start, const double *d_end,
                       size_t nrays, int mode, double *d_out, void *stream);

/* vt_project that also writes every ray's segment count to d_counts (nrays,
 * the counts of vt_project_count) in the same pass; d_counts may be NULL.
 * No reference counterpart: the forward pass of a differentiable projection
 * keeps the counts for vt_adjoint. */
int vt_project_counted(const vt_cloud *cloud, const double *d_start, const double *d_end,
                       size_t nrays, int mode, double *d_out, int64_t *d_counts, void *stream);

/* Projection::makeDetailedProjection (src/projection.cpp:75-144) in two
 * passes.  vt_project_count writes the number of segments of every ray to
 * d_counts (nrays); the caller forms offsets (exclusive scan, nrays + 1) and
 * calls vt_project_segments, which writes the reduced values and, at
 * d_offsets[i], ray i's ordered segments: cell id, entry and exit distance
 * (returnSegments, src/bindings.cpp:206-238). */
int vt_project_count(const vt_cloud *cloud, const double *d_start, const double *d_end,
                     size_t nrays, int64_t *d_counts, void *stream);
int vt_project_segments(const vt_cloud *cloud, const double *d_start, const double *d_end,
                        size_t nrays, int mode, const int64_t *d_offsets, double *d_out,
                        int64_t *d_cell_ids, double *d_s_enter, double *d_s_exit,
                        void *stream);

/* PointCloud::saveTreeToBytes / loadTreeFromBytes (src/pointcloud.cpp:243-
 * 281, bindings 125-132): the host tree as nanoflann 1.3.2's index dump
 * (include/nanoflann.hpp:1023-1071).  save sets *size and writes the bytes
 * when buf has room (cap >= *size).  load checks the dump (leaf size 10, a
 * vind permutation over the cloud's points, leaves tiling the points in
 * pre-order) and replaces the tree; with upload != 0 it also moves it to the
 * current device, as vt_cloud_upload. */
int vt_cloud_save_tree(const vt_cloud *cloud, uint8_t *buf, size_t cap, size_t *size);
int vt_cloud_load_tree(vt_cloud *cloud, const uint8_t *buf, size_t size, int upload);

/* Slice::makeSlice (src/slice.cpp:7-60): the field values of the nearest cell
 * at every pixel of an nx x ny grid at z = depth over extent
 * [xmin, xmax, ymin, ymax], pixel (i, j) at (xmin + dx*i, ymin + dy*j) with
 * dx = (xmax - xmin) / (nx - 1); d_out is (nx * ny) x nfields, pixel
 * (i, j) at row i * ny + j. */
int vt_slice(const vt_cloud *cloud, const double extent[4], size_t nx, size_t ny, double depth,
             double *d_out, void *stream);

/* Declares that the rays of the cloud's coming ray calls are the pixels of
 * row-major images whose rows hold `row` rays (0: no such order, the default).
 * The kernels then take the rays of whole 8-row bands in 8x8-pixel tiles, so a
 * wave's rays are image neighbours that walk the same part of the tree.  Speed
 * only: every output is still written at its ray's own index, with the same
 * bits.  No reference counterpart (the reference's OpenMP loop takes rays in
 * index order, src/projection.cpp:43-68); grid_projection sets it. */
int vt_cloud_set_ray_rows(vt_cloud *cloud, int64_t row);

/* Sets the order in which the kernels take the rays of the cloud's coming
 * calls of exactly n rays: d_order (device, n entries) must be a permutation
 * of [0, n) (checked; VT_EINVAL otherwise); the library keeps a copy.  NULL or
 * n = 0 clears it.  Speed only: every output is still written at its ray's own
 * index, with the same bits.  Fitting loops pass the rays longest first within
 * each ray partition (from the previous pass's segment counts) so the last rays
 * a lane takes are short.  No reference counterpart (src/projection.cpp:43-68
 * takes rays in index order). */
int vt_cloud_set_ray_order(vt_cloud *cloud, const uint32_t *d_order, int64_t n, void *stream);

/* Sets the cloud's ray order for its coming calls of exactly n rays (as
 * vt_cloud_set_ray_order) to longest first, built on the device: fetch slot j
 * names ray j, or with row > 0 the 8x8-tile order of row-major images of
 * `row` rays per row; the slots of each ray partition (vt_tuning.ray_parts
 * contiguous ranges) are cut into blocks of `block` slots (1..32768) and
 * runs of 64, and each block takes its runs by descending key (lower run
 * first among equals).  A run's key is the largest of d_counts (n segment
 * counts of a previous pass over the same rays, device) over its rays, or,
 * when d_counts is NULL, an estimate from the run's middle ray (d_start /
 * d_end, n x 3): its length in box diagonals times the sum of (points per
 * cell)^(1/3) of a 32^3 histogram of the generators over the cloud's box at
 * 32 points along it, x 1024, truncated.  d_order_out (n, optional) receives
 * the order; d_run_keys (optional) the keys of the parts *
 * ceil(ceil(n / parts) / block) * ceil(block / 64) runs.  Speed only.
 * No reference counterpart (src/projection.cpp:43-68 takes rays in index
 * order). */
int vt_cloud_order_rays(vt_cloud *cloud, const double *d_start, const double *d_end,
                        const int64_t *d_counts, int64_t n, int64_t row, int64_t block,
                        uint32_t *d_order_out, int64_t *d_run_keys, void *stream);

/* Replaces the cloud's fields, given on the device in filtered point order
 * (row i belongs to point vt_cloud_orig_ids()[i]; n x nfields, the nfields of
 * vt_cloud_create), without rebuilding the tree, and re-applies the RGBA
 * validity rule of PointCloud::loadPoints (src/pointcloud.cpp:168-179).
 * No reference counterpart: the reference sets fields only in loadPoints
 * (src/pointcloud.cpp:12-191); this is the field update of a fitting loop. */
int vt_cloud_set_fields(vt_cloud *cloud, const double *d_fields, void *stream);

/* Adjoint of vt_project for mode VT_SUM or VT_VOLUME: given d_grad_out
 * (nrays x n_out, the gradient of a scalar loss with respect to the image),
 * writes d_grad_fields (n x nfields, filtered point order), the gradient
 * with respect to the cloud's current fields.  Sum: segment-length-weighted
 * scatter (the derivative of reduce_sum, src/reduction.cpp:8-21).  Volume:
 * reverse compositing per ray (the derivative of reduce_volume_render,
 * src/reduction.cpp:55-83).  d_counts: the rays' segment counts from
 * vt_project_counted, or NULL to count them here.  The segments are those of
 * vt_project_segments;
 * per-cell sums run in ray-major segment order (a deterministic segmented
 * reduction, no floating-point atomics).  No reference counterpart: the
 * reference has no derivatives. */
int vt_adjoint(const vt_cloud *cloud, const double *d_start, const double *d_end, size_t nrays,
               int mode, const int64_t *d_counts, const double *d_grad_out,
               double *d_grad_fields, void *stream);

/* The forward and backward of a differentiable projection with one walk of
 * the rays.  vt_project_recorded is vt_project_counted (modes VT_SUM and
 * VT_VOLUME) that also appends one VT_SEGREC_BYTES record per segment (its
 * cell, its ray and its place along the ray, ds) to d_records, a device pool
 * of `capacity` records (whatever it held before: it need not be cleared),
 * and sets *n_records to the records used, or to (size_t)-1 when the pool
 * was too small (the image and counts are still complete).  vt_adjoint_recorded is vt_adjoint over those records: it
 * scatters them to ray-major order by d_counts and runs the same
 * contributions and per-cell sums, so the two adjoints agree bit for bit.
 * The cloud's fields must be those of the forward pass for VT_VOLUME.
 * No reference counterpart: the reference has no derivatives. */
#define VT_SEGREC_BYTES 32
int vt_project_recorded(const vt_cloud *cloud, const double *d_start, const double *d_end,
                        size_t nrays, int mode, double *d_out, int64_t *d_counts,
                        void *d_records, size_t capacity, size_t *n_records, void *stream);
int vt_adjoint_recorded(const vt_cloud *cloud, size_t nrays, int mode, const int64_t *d_counts,
                        const void *d_records, size_t n_records, const double *d_grad_out,
                        double *d_grad_fields, void *stream);

/* grid.generate_projection_grid (vortrace/grid.py:100-192) on the device,
 * for many views at once: view v's rays follow those of views 0..v-1, ray
 * i * ny + j of a view (pixel x_i, y_j, numpy.linspace pixel centres of
 * grid.py:35-44) starts at (x_i, y_j, bounds[0]) and ends at
 * (x_i, y_j, bounds[1]), rotated about `center` by `rot` (row-major R,
 * grid.py:86-97, computed by the caller) when rotate != 0.  bounds are
 * those after the handedness flip of grid.py:163-172.  d_start / d_end hold
 * sum(nx * ny) x 3 values.  The bits are numpy's for unrotated grids; with a
 * rotation, numpy.dot's rows are evaluated as fma(p2, R_k2, fma(p1, R_k1,
 * p0 * R_k0)), the order of OpenBLAS' x86-64 kernels behind numpy.  No
 * reference counterpart as a C function: the reference builds rays in numpy. */
typedef struct {
  double extent[4]; /* xlo, xhi, ylo, yhi */
  double bounds[2]; /* z of start and end points before the rotation */
  double center[3];
  double rot[9];
  int64_t nx, ny;
  int32_t rotate;
  int32_t pad_;
} vt_view;
int vt_grid_rays(const vt_view *views, size_t nviews, double *d_start, double *d_end,
                 void *stream);

/* Launch parameters of the ray kernels (process-wide): lanes of a wave that
 * run their walk steps together, lanes that run their leaf steps together
 * (both 1..64), ray partitions (1, or 8: one per XCD), and help_events
 * (0..128): lanes with no ray left walk intervals of other lanes' rays of
 * their wave (sum, max, min and count kernels), recording up to this many
 * boundaries of one interval before they hand it back (0: off).  Results do
 * not depend on them, only speed.  No reference counterpart (the reference's
 * OpenMP loop has schedule(dynamic,256) fixed, src/projection.cpp:43);
 * vortrace_amd.tuning.autotune picks them. */
typedef struct {
  int walk_batch, leaf_batch, ray_parts, help_events;
} vt_tuning;
int vt_set_tuning(const vt_tuning *t);
int vt_get_tuning(vt_tuning *t);

/* Resident lanes of the forward kernel a vt_project_counted (recorded = 0)
 * or vt_project_recorded (recorded != 0) call of `mode` on this cloud
 * launches: its persistent grid times the workgroup size.  A launch with
 * fewer rays leaves lanes idle; a recording launch leaves at most one partly
 * filled chunk of 32 records per lane.  No reference counterpart (the
 * reference's OpenMP loop has no fixed lane count, src/projection.cpp:43). */
int vt_cloud_lanes(const vt_cloud *cloud, int mode, int recorded, int64_t *lanes);

/* Per-call statistics of the last vt_project* call on this thread: total
 * segments (cell crossings), kNN queries run, the deepest walk and kNN
 * stacks, K_EXPAND searches, rays re-run by the K_EXPAND-capable kernel, the
 * kernels' time, and the intervals of rays walked by other lanes than the
 * ray's own (vt_tuning.help_events) and of those the ones handed back
 * unwalked. */
typedef struct {
  int64_t segments, queries, max_walk_depth, max_knn_depth, slow_queries, deferred_rays;
  float kernel_ms;
  int64_t intervals, intervals_back;
} vt_call_stats;
int vt_last_call_stats(vt_call_stats *out);

#ifdef __cplusplus
}
#endif
#endif
