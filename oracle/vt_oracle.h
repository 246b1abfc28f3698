/* CPU restatement of the reference's ray-projection path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  The product (vortrace_amd/)
 * never calls it.
 *
 * It restates, in plain C with OpenMP over rays:
 *   PointCloud::loadPoints / buildTree / queryTreeK   (src/pointcloud.cpp)
 *   nanoflann 1.3.2 build + kNN search                 (include/nanoflann.hpp)
 *   Ray::walk and helpers                              (src/ray.cpp)
 *   reduce_sum / max / min / volume_render             (src/reduction.cpp)
 *   Projection::makeProjection / makeDetailedProjection (src/projection.cpp)
 */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vto_cloud vto_cloud;

/* Return codes: 0 ok, 1 invalid argument (ValueError), 2 runtime error. */
const char *vto_last_error(void);

/* Test-only: plant mistake number m (0 = none) to show the pins catch it. */
void vto_set_mutation(int m);

int vto_cloud_create(const double *pos, size_t npart, const double *fields, size_t nfields,
                     const double subbox[6], const double *vol, size_t nvol, int periodic,
                     int build_tree, vto_cloud **out);
void vto_cloud_destroy(vto_cloud *c);
size_t vto_cloud_npart(const vto_cloud *c);
int vto_cloud_valid_rgba(const vto_cloud *c);
double vto_cloud_pad(const vto_cloud *c);
void vto_cloud_orig_ids(const vto_cloud *c, int64_t *out);
int vto_cloud_num_warnings(const vto_cloud *c);
const char *vto_cloud_warning(const vto_cloud *c, int i);

/* Tree export in pre-order: feat = -1 for a leaf (a = left, b = right into
 * vind), else the cut dimension with divlow/divhigh.  vind has npart entries. */
size_t vto_tree_nnodes(const vto_cloud *c);
void vto_tree_export(const vto_cloud *c, int64_t *vind, int32_t *feat, int64_t *a, int64_t *b,
                     double *divlow, double *divhigh, double *root_bbox6);

int vto_query(const vto_cloud *c, const double *q, size_t nq, size_t k, int64_t *ids,
              double *r2);

/* mode: 0 sum, 1 max, 2 min, 3 volume.  out: nrays x n_out. nthreads <= 0
 * leaves OpenMP's default. */
int vto_project(const vto_cloud *c, const double *start, const double *end, size_t nrays,
                int mode, double *out, int nthreads);

typedef struct vto_segments vto_segments;
int vto_detailed(const vto_cloud *c, const double *start, const double *end, size_t nrays,
                 int mode, int nthreads, vto_segments **out, size_t *nseg);
void vto_detailed_copy(const vto_segments *s, double *values, int64_t *cell_ids,
                       double *s_enter, double *s_exit, int64_t *offsets);
void vto_detailed_destroy(vto_segments *s);

/* p (n x 3) rotated about c by R (row-major) in numpy.dot's evaluation order. */
void vto_rotate_about(const double *p, size_t n, const double *R, const double *c, double *out);

#ifdef __cplusplus
}
#endif
#endif
