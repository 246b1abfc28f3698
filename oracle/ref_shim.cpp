// C-ABI shim over the reference's own C++ core, for use as a checker only.
//
// This file is compiled together with the reference sources as they lie in the
// reference checkout (src/pointcloud.cpp, src/ray.cpp, src/reduction.cpp,
// src/projection.cpp, include/nanoflann.hpp) by oracle/Makefile into
// oracle/_ref/libvtref.so.  Nothing here is product code: tests and the CPU
// baseline leg of bench.py load it through ctypes to compare against.
//
// Entry points mirror the pybind11 module the reference exposes
// (src/bindings.cpp:109-238): PointCloud.loadPoints/buildTree/queryTree,
// Projection.makeProjection/makeDetailedProjection/returnSegments and the
// nanoflann tree dump behind saveTreeToBytes (src/pointcloud.cpp:242-261).

#include "pointcloud.hpp"
#include "projection.hpp"
#include "ray.hpp"
#include "reduction.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;
thread_local int g_err_kind = 0;  // 1 = invalid_argument, 2 = other
std::vector<std::string> g_warnings;

int fail(const std::exception &e, int kind) {
  g_err = e.what();
  g_err_kind = kind;
  return kind;
}

#define GUARD_BEGIN try {
#define GUARD_END                                                     \
  }                                                                   \
  catch (const std::invalid_argument &e) { return fail(e, 1); }       \
  catch (const std::exception &e) { return fail(e, 2); }              \
  return 0;

struct RefProjection {
  Projection proj;
  RefProjection(const double *s, const double *e, size_t n) : proj(s, e, n) {}
};

}  // namespace

extern "C" {

const char *vtref_last_error(void) { return g_err.c_str(); }

// Warnings raised by the reference through vortrace::warn are collected so the
// caller can compare them with the drop-in's warning texts.
void vtref_capture_warnings(void) {
  vortrace::warning_handler = [](const std::string &m) { g_warnings.push_back(m); };
}
size_t vtref_num_warnings(void) { return g_warnings.size(); }
const char *vtref_warning(size_t i) { return i < g_warnings.size() ? g_warnings[i].c_str() : ""; }
void vtref_clear_warnings(void) { g_warnings.clear(); }

int vtref_cloud_create(const double *pos, size_t npart, const double *fields,
                       size_t nfields, const double *subbox6, const double *vol,
                       size_t nvol, int periodic, int build_tree, void **out) {
  GUARD_BEGIN
  auto *pc = new PointCloud();
  std::array<Float, 6> sb;
  for (int i = 0; i < 6; i++) sb[i] = subbox6[i];
  try {
    pc->loadPoints(pos, npart, fields, npart, nfields, sb, nvol ? vol : nullptr, nvol,
                   periodic != 0);
    if (build_tree) pc->buildTree();
  } catch (...) {
    delete pc;
    throw;
  }
  *out = pc;
  GUARD_END
}

void vtref_cloud_destroy(void *h) { delete static_cast<PointCloud *>(h); }

size_t vtref_cloud_npart(void *h) { return static_cast<PointCloud *>(h)->get_npart(); }

int vtref_cloud_valid_rgba(void *h) { return static_cast<PointCloud *>(h)->get_valid_rgba(); }

double vtref_cloud_pad(void *h) { return static_cast<PointCloud *>(h)->get_pad(); }

void vtref_cloud_orig_ids(void *h, int64_t *out) {
  const auto &ids = static_cast<PointCloud *>(h)->get_orig_ids();
  for (size_t i = 0; i < ids.size(); i++) out[i] = static_cast<int64_t>(ids[i]);
}

// k-NN for nq query points; ids/r2 are nq*k, row-major.
int vtref_query(void *h, const double *q, size_t nq, size_t k, int64_t *ids, double *r2) {
  GUARD_BEGIN
  auto *pc = static_cast<PointCloud *>(h);
  std::vector<size_t> tmp(k);
  for (size_t i = 0; i < nq; i++) {
    Point p = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    pc->queryTreeK(p, k, tmp.data(), r2 + i * k);
    for (size_t j = 0; j < k; j++) ids[i * k + j] = static_cast<int64_t>(tmp[j]);
  }
  GUARD_END
}

// Serialized nanoflann index (the bytes behind saveTreeToBytes).
int vtref_tree_bytes(void *h, char **buf, size_t *len) {
  GUARD_BEGIN
  auto v = static_cast<PointCloud *>(h)->saveTreeToBuffer();
  *len = v.size();
  *buf = static_cast<char *>(malloc(v.size()));
  std::memcpy(*buf, v.data(), v.size());
  GUARD_END
}
void vtref_free(void *p) { free(p); }

// Projection::makeProjection; out has nrays * n_out entries.
int vtref_project(void *h, const double *start, const double *end, size_t nrays, int mode,
                  double *out) {
  GUARD_BEGIN
  Projection proj(start, end, nrays);
  proj.makeProjection(*static_cast<PointCloud *>(h), static_cast<ReductionMode>(mode));
  const auto &d = proj.getProjectionData();
  std::memcpy(out, d.data(), d.size() * sizeof(double));
  GUARD_END
}

// Projection::makeDetailedProjection; returns a handle holding the segments.
int vtref_detailed(void *h, const double *start, const double *end, size_t nrays, int mode,
                   void **out, size_t *nseg) {
  GUARD_BEGIN
  auto *rp = new RefProjection(start, end, nrays);
  try {
    rp->proj.makeDetailedProjection(*static_cast<PointCloud *>(h),
                                    static_cast<ReductionMode>(mode));
  } catch (...) {
    delete rp;
    throw;
  }
  *out = rp;
  *nseg = rp->proj.getSegCellIds().size();
  GUARD_END
}

void vtref_detailed_copy(void *hp, double *values, int64_t *cell_ids, double *s_enter,
                         double *s_exit, int64_t *offsets) {
  auto &p = static_cast<RefProjection *>(hp)->proj;
  const auto &d = p.getProjectionData();
  std::memcpy(values, d.data(), d.size() * sizeof(double));
  const auto &ids = p.getSegCellIds();
  for (size_t i = 0; i < ids.size(); i++) cell_ids[i] = static_cast<int64_t>(ids[i]);
  std::memcpy(s_enter, p.getSegSEnter().data(), ids.size() * sizeof(double));
  std::memcpy(s_exit, p.getSegSExit().data(), ids.size() * sizeof(double));
  const auto &off = p.getSegOffsets();
  for (size_t i = 0; i < off.size(); i++) offsets[i] = static_cast<int64_t>(off[i]);
}

void vtref_detailed_destroy(void *hp) { delete static_cast<RefProjection *>(hp); }

}  // extern "C"
