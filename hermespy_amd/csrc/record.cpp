// HDF5 drop store (declared in include/hermespy_amd_h5.h): host code over the libhdf5 C API.
#include "hermespy_amd_h5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

namespace {

constexpr int kOk = 0;
constexpr int kErrInvalid = -1;  // HPX_ERR_INVALID
constexpr int kErrIo = -4;

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

// Closes an HDF5 identifier at scope exit.
struct Id {
  hid_t id;
  herr_t (*close)(hid_t);
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  operator hid_t() const { return id; }
  bool ok() const { return id >= 0; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

// memory type of a dtype code (a new identifier the caller closes), or -1
hid_t mem_type(int dtype) {
  switch (dtype) {
    case HPX_H5_U8: return H5Tcopy(H5T_NATIVE_UINT8);
    case HPX_H5_I8: return H5Tcopy(H5T_NATIVE_INT8);
    case HPX_H5_I32: return H5Tcopy(H5T_NATIVE_INT32);
    case HPX_H5_U32: return H5Tcopy(H5T_NATIVE_UINT32);
    case HPX_H5_I64: return H5Tcopy(H5T_NATIVE_INT64);
    case HPX_H5_F32: return H5Tcopy(H5T_NATIVE_FLOAT);
    case HPX_H5_F64: return H5Tcopy(H5T_NATIVE_DOUBLE);
    case HPX_H5_C64:
    case HPX_H5_C128: {
      const bool dbl = dtype == HPX_H5_C128;
      const size_t part = dbl ? sizeof(double) : sizeof(float);
      const hid_t t = H5Tcreate(H5T_COMPOUND, 2 * part);
      if (t < 0) return -1;
      H5Tinsert(t, "r", 0, dbl ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT);
      H5Tinsert(t, "i", part, dbl ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT);
      return t;
    }
    case HPX_H5_STR: {
      const hid_t t = H5Tcopy(H5T_C_S1);
      if (t < 0) return -1;
      H5Tset_size(t, H5T_VARIABLE);
      H5Tset_cset(t, H5T_CSET_UTF8);
      return t;
    }
  }
  return -1;
}

size_t type_size(int dtype) {
  switch (dtype) {
    case HPX_H5_U8:
    case HPX_H5_I8: return 1;
    case HPX_H5_I32:
    case HPX_H5_U32:
    case HPX_H5_F32: return 4;
    case HPX_H5_I64:
    case HPX_H5_F64:
    case HPX_H5_C64: return 8;
    case HPX_H5_C128: return 16;
  }
  return 0;
}

// dtype code of a stored type, or -1
int code_of(hid_t t) {
  const H5T_class_t c = H5Tget_class(t);
  const size_t size = H5Tget_size(t);
  if (c == H5T_INTEGER) {
    const bool sgn = H5Tget_sign(t) == H5T_SGN_2;
    if (size == 1) return sgn ? HPX_H5_I8 : HPX_H5_U8;
    if (size == 4) return sgn ? HPX_H5_I32 : HPX_H5_U32;
    if (size == 8) return HPX_H5_I64;
  } else if (c == H5T_FLOAT) {
    if (size == 4) return HPX_H5_F32;
    if (size == 8) return HPX_H5_F64;
  } else if (c == H5T_COMPOUND && H5Tget_nmembers(t) == 2) {
    if (size == 8) return HPX_H5_C64;
    if (size == 16) return HPX_H5_C128;
  } else if (c == H5T_STRING) {
    return HPX_H5_STR;
  }
  return -1;
}

hid_t link_create_plist() {
  const hid_t p = H5Pcreate(H5P_LINK_CREATE);
  if (p >= 0) H5Pset_create_intermediate_group(p, 1);
  return p;
}

bool valid_path(const char* p) { return p && p[0]; }

}  // namespace

extern "C" int hpx_h5_last_error(char* buffer, size_t size) {
  if (!buffer || size == 0) return kErrInvalid;
  snprintf(buffer, size, "%s", g_error.c_str());
  return kOk;
}

extern "C" int hpx_h5_open(const char* path, int mode, int64_t* handle) {
  quiet();
  if (!valid_path(path) || !handle) return fail(kErrInvalid, "hpx_h5_open: null path or handle");
  hid_t f = -1;
  switch (mode) {
    case HPX_H5_READ: f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT); break;
    case HPX_H5_APPEND: f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT); break;
    case HPX_H5_CREATE: f = H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT); break;
    case HPX_H5_TRUNCATE: f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT); break;
    default: return fail(kErrInvalid, "hpx_h5_open: unknown mode %d", mode);
  }
  if (f < 0) return fail(kErrIo, "hpx_h5_open: cannot open '%s' (mode %d)", path, mode);
  *handle = static_cast<int64_t>(f);
  return kOk;
}

extern "C" int hpx_h5_close(int64_t handle) {
  quiet();
  if (H5Fclose(static_cast<hid_t>(handle)) < 0) return fail(kErrIo, "hpx_h5_close: invalid handle");
  return kOk;
}

extern "C" int hpx_h5_exists(int64_t handle, const char* path) {
  quiet();
  if (!valid_path(path)) return fail(kErrInvalid, "hpx_h5_exists: null path");
  if (!strcmp(path, "/")) return 1;
  // check every prefix: H5Lexists fails on a missing intermediate group
  std::string p(path);
  size_t pos = p[0] == '/' ? 1 : 0;
  while (true) {
    const size_t next = p.find('/', pos);
    const std::string prefix = p.substr(0, next);
    const htri_t e = H5Lexists(static_cast<hid_t>(handle), prefix.c_str(), H5P_DEFAULT);
    if (e < 0) return fail(kErrIo, "hpx_h5_exists: cannot query '%s'", prefix.c_str());
    if (e == 0) return 0;
    if (next == std::string::npos) return 1;
    pos = next + 1;
  }
}

extern "C" int hpx_h5_create_group(int64_t handle, const char* path) {
  quiet();
  const int e = hpx_h5_exists(handle, path);
  if (e < 0) return e;
  if (e == 1) return kOk;
  Id lcpl(link_create_plist(), H5Pclose);
  Id g(H5Gcreate2(static_cast<hid_t>(handle), path, lcpl, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (!g.ok()) return fail(kErrIo, "hpx_h5_create_group: cannot create '%s'", path);
  return kOk;
}

extern "C" int hpx_h5_set_attr(int64_t handle, const char* object, const char* name, int dtype, int64_t count,
                               const void* data) {
  quiet();
  if (!valid_path(object) || !name || !data) return fail(kErrInvalid, "hpx_h5_set_attr: null argument");
  if (dtype != HPX_H5_STR && count < 1) return fail(kErrInvalid, "hpx_h5_set_attr: count must be positive");
  const hid_t f = static_cast<hid_t>(handle);
  Id t(mem_type(dtype), H5Tclose);
  if (!t.ok()) return fail(kErrInvalid, "hpx_h5_set_attr: unknown element type %d", dtype);
  const htri_t e = H5Aexists_by_name(f, object, name, H5P_DEFAULT);
  if (e < 0) return fail(kErrIo, "hpx_h5_set_attr: no object '%s'", object);
  if (e > 0 && H5Adelete_by_name(f, object, name, H5P_DEFAULT) < 0)
    return fail(kErrIo, "hpx_h5_set_attr: cannot replace '%s' of '%s'", name, object);
  hsize_t dims[1] = {static_cast<hsize_t>(count)};
  Id space(dtype == HPX_H5_STR || count == 1 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, dims, nullptr), H5Sclose);
  Id a(H5Acreate_by_name(f, object, name, t, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return fail(kErrIo, "hpx_h5_set_attr: cannot create '%s' on '%s'", name, object);
  herr_t st;
  if (dtype == HPX_H5_STR) {
    const char* s = static_cast<const char*>(data);
    st = H5Awrite(a, t, &s);
  } else {
    st = H5Awrite(a, t, data);
  }
  if (st < 0) return fail(kErrIo, "hpx_h5_set_attr: cannot write '%s'", name);
  return kOk;
}

extern "C" int hpx_h5_get_attr(int64_t handle, const char* object, const char* name, int dtype, int64_t capacity,
                               void* data, int64_t* count) {
  quiet();
  if (!valid_path(object) || !name || !data || !count) return fail(kErrInvalid, "hpx_h5_get_attr: null argument");
  Id a(H5Aopen_by_name(static_cast<hid_t>(handle), object, name, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok()) return fail(kErrIo, "hpx_h5_get_attr: no attribute '%s' on '%s'", name, object);
  Id space(H5Aget_space(a), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(space);
  Id ft(H5Aget_type(a), H5Tclose);
  if (dtype == HPX_H5_STR) {
    if (H5Tget_class(ft) != H5T_STRING || n != 1) return fail(kErrInvalid, "hpx_h5_get_attr: '%s' is no string", name);
    std::string value;
    if (H5Tis_variable_str(ft) > 0) {
      Id t(mem_type(HPX_H5_STR), H5Tclose);
      char* s = nullptr;
      if (H5Aread(a, t, &s) < 0) return fail(kErrIo, "hpx_h5_get_attr: cannot read '%s'", name);
      value = s ? s : "";
      H5free_memory(s);
    } else {
      const size_t len = H5Tget_size(ft);
      std::vector<char> buf(len + 1, 0);
      Id t(H5Tcopy(ft), H5Tclose);
      if (H5Aread(a, t, buf.data()) < 0) return fail(kErrIo, "hpx_h5_get_attr: cannot read '%s'", name);
      value.assign(buf.data(), strnlen(buf.data(), len));
    }
    *count = static_cast<int64_t>(value.size());
    if (static_cast<int64_t>(value.size()) + 1 > capacity) return fail(kErrInvalid, "hpx_h5_get_attr: buffer too small");
    memcpy(data, value.c_str(), value.size() + 1);
    return kOk;
  }
  *count = static_cast<int64_t>(n);
  if (n > capacity) return fail(kErrInvalid, "hpx_h5_get_attr: '%s' holds %lld elements", name, (long long)n);
  Id t(mem_type(dtype), H5Tclose);
  if (!t.ok()) return fail(kErrInvalid, "hpx_h5_get_attr: unknown element type %d", dtype);
  if (H5Aread(a, t, data) < 0) return fail(kErrIo, "hpx_h5_get_attr: cannot convert '%s'", name);
  return kOk;
}

extern "C" int hpx_h5_append(int64_t handle, const char* path, int dtype, int rank, const int64_t* row_shape,
                             int64_t rows, const void* data) {
  quiet();
  if (!valid_path(path) || (rows > 0 && !data) || rank < 0 || rank >= HPX_H5_MAX_RANK || (rank > 0 && !row_shape))
    return fail(kErrInvalid, "hpx_h5_append: invalid arguments");
  if (rows < 0) return fail(kErrInvalid, "hpx_h5_append: negative row count");
  if (dtype == HPX_H5_STR || type_size(dtype) == 0) return fail(kErrInvalid, "hpx_h5_append: unsupported type %d", dtype);
  const hid_t f = static_cast<hid_t>(handle);
  Id t(mem_type(dtype), H5Tclose);
  hsize_t dims[HPX_H5_MAX_RANK], maxd[HPX_H5_MAX_RANK], chunk[HPX_H5_MAX_RANK];
  size_t row_bytes = type_size(dtype);
  for (int i = 0; i < rank; ++i) {
    if (row_shape[i] < 1) return fail(kErrInvalid, "hpx_h5_append: row shape must be positive");
    dims[i + 1] = maxd[i + 1] = chunk[i + 1] = static_cast<hsize_t>(row_shape[i]);
    row_bytes *= static_cast<size_t>(row_shape[i]);
  }
  const int e = hpx_h5_exists(handle, path);
  if (e < 0) return e;
  if (e == 0) {
    dims[0] = 0;
    maxd[0] = H5S_UNLIMITED;
    chunk[0] = static_cast<hsize_t>(row_bytes >= (1u << 20) ? 1 : (1u << 20) / row_bytes);
    Id space(H5Screate_simple(rank + 1, dims, maxd), H5Sclose);
    Id dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    H5Pset_chunk(dcpl, rank + 1, chunk);
    Id lcpl(link_create_plist(), H5Pclose);
    Id d(H5Dcreate2(f, path, t, space, lcpl, dcpl, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(kErrIo, "hpx_h5_append: cannot create '%s'", path);
  }
  Id d(H5Dopen2(f, path, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(kErrIo, "hpx_h5_append: cannot open '%s'", path);
  Id space(H5Dget_space(d), H5Sclose);
  hsize_t cur[HPX_H5_MAX_RANK];
  if (H5Sget_simple_extent_ndims(space) != rank + 1) return fail(kErrInvalid, "hpx_h5_append: '%s' has another rank", path);
  H5Sget_simple_extent_dims(space, cur, nullptr);
  for (int i = 1; i <= rank; ++i)
    if (cur[i] != dims[i]) return fail(kErrInvalid, "hpx_h5_append: '%s' has another row shape", path);
  {
    Id ft(H5Dget_type(d), H5Tclose);
    if (code_of(ft) != dtype) return fail(kErrInvalid, "hpx_h5_append: '%s' has another element type", path);
  }
  if (rows == 0) return kOk;
  hsize_t grown[HPX_H5_MAX_RANK];
  for (int i = 0; i <= rank; ++i) grown[i] = cur[i];
  grown[0] = cur[0] + static_cast<hsize_t>(rows);
  if (H5Dset_extent(d, grown) < 0) return fail(kErrIo, "hpx_h5_append: cannot extend '%s'", path);
  Id fspace(H5Dget_space(d), H5Sclose);
  hsize_t start[HPX_H5_MAX_RANK] = {0}, count[HPX_H5_MAX_RANK];
  start[0] = cur[0];
  count[0] = static_cast<hsize_t>(rows);
  for (int i = 1; i <= rank; ++i) count[i] = dims[i];
  H5Sselect_hyperslab(fspace, H5S_SELECT_SET, start, nullptr, count, nullptr);
  Id mspace(H5Screate_simple(rank + 1, count, nullptr), H5Sclose);
  if (H5Dwrite(d, t, mspace, fspace, H5P_DEFAULT, data) < 0) return fail(kErrIo, "hpx_h5_append: cannot write '%s'", path);
  return kOk;
}

extern "C" int hpx_h5_shape(int64_t handle, const char* path, int* dtype, int* rank, int64_t* shape) {
  quiet();
  if (!valid_path(path) || !dtype || !rank || !shape) return fail(kErrInvalid, "hpx_h5_shape: null argument");
  Id d(H5Dopen2(static_cast<hid_t>(handle), path, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(kErrIo, "hpx_h5_shape: no dataset '%s'", path);
  Id ft(H5Dget_type(d), H5Tclose);
  Id space(H5Dget_space(d), H5Sclose);
  const int r = H5Sget_simple_extent_ndims(space);
  if (r < 0 || r > HPX_H5_MAX_RANK) return fail(kErrInvalid, "hpx_h5_shape: '%s' has rank %d", path, r);
  hsize_t dims[HPX_H5_MAX_RANK];
  H5Sget_simple_extent_dims(space, dims, nullptr);
  *dtype = code_of(ft);
  *rank = r;
  for (int i = 0; i < r; ++i) shape[i] = static_cast<int64_t>(dims[i]);
  return kOk;
}

extern "C" int hpx_h5_read(int64_t handle, const char* path, int64_t row0, int64_t rows, int dtype, void* data) {
  quiet();
  if (!valid_path(path) || (rows > 0 && !data) || rows < 0 || row0 < 0)
    return fail(kErrInvalid, "hpx_h5_read: invalid arguments");
  Id t(mem_type(dtype), H5Tclose);
  if (!t.ok() || dtype == HPX_H5_STR) return fail(kErrInvalid, "hpx_h5_read: unsupported type %d", dtype);
  Id d(H5Dopen2(static_cast<hid_t>(handle), path, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(kErrIo, "hpx_h5_read: no dataset '%s'", path);
  Id fspace(H5Dget_space(d), H5Sclose);
  const int r = H5Sget_simple_extent_ndims(fspace);
  hsize_t dims[HPX_H5_MAX_RANK];
  H5Sget_simple_extent_dims(fspace, dims, nullptr);
  if (r == 0) {  // scalar dataset
    if (row0 != 0 || rows != 1) return fail(kErrInvalid, "hpx_h5_read: '%s' is a scalar", path);
    if (H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return fail(kErrIo, "hpx_h5_read: cannot read '%s'", path);
    return kOk;
  }
  if (static_cast<hsize_t>(row0 + rows) > dims[0]) return fail(kErrInvalid, "hpx_h5_read: rows beyond '%s'", path);
  if (rows == 0) return kOk;
  hsize_t start[HPX_H5_MAX_RANK] = {0}, count[HPX_H5_MAX_RANK];
  start[0] = static_cast<hsize_t>(row0);
  count[0] = static_cast<hsize_t>(rows);
  for (int i = 1; i < r; ++i) count[i] = dims[i];
  H5Sselect_hyperslab(fspace, H5S_SELECT_SET, start, nullptr, count, nullptr);
  Id mspace(H5Screate_simple(r, count, nullptr), H5Sclose);
  if (H5Dread(d, t, mspace, fspace, H5P_DEFAULT, data) < 0) return fail(kErrIo, "hpx_h5_read: cannot convert '%s'", path);
  return kOk;
}

namespace {
herr_t collect(hid_t, const char* name, const H5L_info_t*, void* op) {
  auto* out = static_cast<std::string*>(op);
  out->append(name);
  out->push_back('\n');
  return 0;
}
}  // namespace

extern "C" int hpx_h5_list(int64_t handle, const char* path, char* names, int64_t capacity, int64_t* size) {
  quiet();
  if (!valid_path(path) || !size) return fail(kErrInvalid, "hpx_h5_list: null argument");
  Id g(H5Gopen2(static_cast<hid_t>(handle), path, H5P_DEFAULT), H5Gclose);
  if (!g.ok()) return fail(kErrIo, "hpx_h5_list: no group '%s'", path);
  std::string all;
  hsize_t idx = 0;
  if (H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, collect, &all) < 0)
    return fail(kErrIo, "hpx_h5_list: cannot iterate '%s'", path);
  *size = static_cast<int64_t>(all.size() + 1);
  if (!names || capacity < *size) return fail(kErrInvalid, "hpx_h5_list: buffer of %lld bytes needed", (long long)*size);
  memcpy(names, all.c_str(), all.size() + 1);
  return kOk;
}
