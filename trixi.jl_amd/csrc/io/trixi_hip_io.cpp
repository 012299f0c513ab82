// Solution, restart and TreeMesh mesh files in the reference's HDF5 layout
// (host-only; linked against the system's libhdf5). The C ABI and the
// reference functions each entry point mirrors are in include/trixi_hip_io.h.
#include "trixi_hip_io.h"

#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// One HDF5 identifier, closed with its own close function.
struct Id {
  hid_t id;
  herr_t (*close)(hid_t);
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Id() {
    if (id >= 0) close(id);
  }
  Id(const Id &) = delete;
  Id &operator=(const Id &) = delete;
  bool ok() const { return id >= 0; }
};

// HDF5 prints its error stack by default; the library reports through its own
// codes instead, so printing is off for the duration of one call.
struct Quiet {
  H5E_auto2_t fn = nullptr;
  void *data = nullptr;
  Quiet() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Quiet() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

// fixed-length UTF-8 string of exactly the string's bytes (at least 1), as HDF5.jl writes a String
hid_t string_type(size_t len) {
  hid_t t = H5Tcopy(H5T_C_S1);
  if (t < 0) return t;
  H5Tset_size(t, len ? len : 1);
  H5Tset_cset(t, H5T_CSET_UTF8);
  H5Tset_strpad(t, H5T_STR_NULLTERM);
  return t;
}

bool put_attr(hid_t obj, const char *name, hid_t ftype, hid_t mtype, int rank, const hsize_t *dims,
              const void *data) {
  Id sp(rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, dims, nullptr), H5Sclose);
  if (!sp.ok()) return false;
  Id a(H5Acreate2(obj, name, ftype, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  return a.ok() && H5Awrite(a.id, mtype, data) >= 0;
}

bool put_i64(hid_t obj, const char *name, int64_t v) {
  return put_attr(obj, name, H5T_STD_I64LE, H5T_NATIVE_INT64, 0, nullptr, &v);
}

bool put_f64(hid_t obj, const char *name, double v) {
  return put_attr(obj, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 0, nullptr, &v);
}

bool put_str(hid_t obj, const char *name, const char *s) {
  const size_t n = std::strlen(s);
  Id t(string_type(n), H5Tclose);
  if (!t.ok()) return false;
  std::vector<char> buf(n ? n : 1, '\0');
  std::memcpy(buf.data(), s, n);
  return put_attr(obj, name, t.id, t.id, 0, nullptr, buf.data());
}

bool put_dataset(hid_t file, const char *name, hid_t ftype, hid_t mtype, int rank, const hsize_t *dims,
                 const void *data, const char *attr_name = nullptr) {
  Id sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  if (!sp.ok()) return false;
  Id ds(H5Dcreate2(file, name, ftype, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!ds.ok() || H5Dwrite(ds.id, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return false;
  return !attr_name || put_str(ds.id, "name", attr_name);
}

bool attr_points(hid_t a, hssize_t want) {
  Id sp(H5Aget_space(a), H5Sclose);
  return sp.ok() && H5Sget_simple_extent_npoints(sp.id) == want;
}

bool get_i64(hid_t obj, const char *name, int64_t *out) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  return a.ok() && attr_points(a.id, 1) && H5Aread(a.id, H5T_NATIVE_INT64, out) >= 0;
}

bool get_f64(hid_t obj, const char *name, double *out) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  return a.ok() && attr_points(a.id, 1) && H5Aread(a.id, H5T_NATIVE_DOUBLE, out) >= 0;
}

bool get_f64_vec(hid_t obj, const char *name, double *out, hssize_t n) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  return a.ok() && attr_points(a.id, n) && H5Aread(a.id, H5T_NATIVE_DOUBLE, out) >= 0;
}

// Bools: an 8-bit bitfield (as written here), or an 8-bit integer / enum
bool get_bool_vec(hid_t obj, const char *name, uint8_t *out, hssize_t n) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || !attr_points(a.id, n)) return false;
  Id t(H5Aget_type(a.id), H5Tclose);
  if (!t.ok() || H5Tget_size(t.id) != 1) return false;
  const H5T_class_t cls = H5Tget_class(t.id);
  hid_t mtype = cls == H5T_BITFIELD ? H5T_NATIVE_B8 : cls == H5T_INTEGER ? H5T_NATIVE_UINT8 : -1;
  if (cls == H5T_ENUM) return H5Aread(a.id, t.id, out) >= 0;
  return mtype >= 0 && H5Aread(a.id, mtype, out) >= 0;
}

// string attribute, fixed or variable length, trailing NULs and spaces removed
bool get_str(hid_t obj, const char *name, std::string *out) {
  Id a(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || !attr_points(a.id, 1)) return false;
  Id t(H5Aget_type(a.id), H5Tclose);
  if (!t.ok() || H5Tget_class(t.id) != H5T_STRING) return false;
  if (H5Tis_variable_str(t.id) > 0) {
    Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mt.id, H5T_VARIABLE);
    char *p = nullptr;
    if (H5Aread(a.id, mt.id, &p) < 0) return false;
    *out = p ? p : "";
    H5free_memory(p);
  } else {
    const size_t n = H5Tget_size(t.id);
    std::vector<char> buf(n + 1, '\0');
    if (H5Aread(a.id, t.id, buf.data()) < 0) return false;
    out->assign(buf.data(), strnlen(buf.data(), n));
  }
  while (!out->empty() && out->back() == ' ') out->pop_back();
  return true;
}

bool copy_str(const std::string &s, char *dst, size_t cap) {
  if (s.size() + 1 > cap) return false;
  std::memcpy(dst, s.c_str(), s.size() + 1);
  return true;
}

bool read_dataset(hid_t file, const char *name, hid_t mtype, hssize_t want, void *out) {
  Id ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return false;
  Id sp(H5Dget_space(ds.id), H5Sclose);
  if (!sp.ok() || H5Sget_simple_extent_npoints(sp.id) != want) return false;
  return H5Dread(ds.id, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0;
}

bool valid_header(const trixi_io_header *h) {
  return h && (h->ndims == 2 || h->ndims == 3) && h->polydeg >= 1 && h->n_vars >= 1 && h->n_elements >= 1 &&
         strnlen(h->equations, sizeof(h->equations)) < sizeof(h->equations) &&
         strnlen(h->mesh_type, sizeof(h->mesh_type)) < sizeof(h->mesh_type) &&
         strnlen(h->mesh_file, sizeof(h->mesh_file)) < sizeof(h->mesh_file);
}

}  // namespace

extern "C" {

const char *trixi_io_last_error(void) { return g_err.c_str(); }

int trixi_io_write_solution(const char *filename, const trixi_io_header *h, int64_t n_nodes, const double *u,
                            const char *const *varnames, int n_element_variables, const double *element_values,
                            const char *const *element_names) {
  if (!filename || !valid_header(h) || n_nodes < 1 || !u || !varnames || n_element_variables < 0 ||
      (n_element_variables > 0 && (!element_values || !element_names)))
    return fail(TRIXI_IO_EINVAL, "invalid arguments");
  Quiet quiet;
  Id f(H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(TRIXI_IO_EFILE, std::string("cannot create ") + filename);
  // attribute order of save_solution_file / save_restart_file
  bool ok = put_i64(f.id, "ndims", h->ndims) && put_str(f.id, "equations", h->equations) &&
            put_i64(f.id, "polydeg", h->polydeg) && put_i64(f.id, "n_vars", h->n_vars) &&
            put_i64(f.id, "n_elements", h->n_elements) && put_str(f.id, "mesh_type", h->mesh_type) &&
            put_str(f.id, "mesh_file", h->mesh_file) && put_f64(f.id, "time", h->time) &&
            put_f64(f.id, "dt", h->dt) && put_i64(f.id, "timestep", h->timestep);
  const int64_t nv = h->n_vars;
  const hsize_t count = (hsize_t)n_nodes * (hsize_t)h->n_elements;
  std::vector<double> column(count);
  for (int64_t v = 0; ok && v < nv; ++v) {
    if (!varnames[v]) return fail(TRIXI_IO_EINVAL, "missing variable name");
    for (hsize_t k = 0; k < count; ++k) column[k] = u[(size_t)v + (size_t)nv * k];
    const std::string name = "variables_" + std::to_string(v + 1);
    ok = put_dataset(f.id, name.c_str(), H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 1, &count, column.data(), varnames[v]);
  }
  const hsize_t ne = (hsize_t)h->n_elements;
  for (int k = 0; ok && k < n_element_variables; ++k) {
    if (!element_names[k]) return fail(TRIXI_IO_EINVAL, "missing element variable name");
    const std::string name = "element_variables_" + std::to_string(k + 1);
    ok = put_dataset(f.id, name.c_str(), H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 1, &ne, element_values + (size_t)k * ne,
                     element_names[k]);
  }
  if (!ok) return fail(TRIXI_IO_EFILE, std::string("writing ") + filename + " failed");
  return TRIXI_IO_OK;
}

int trixi_io_read_header(const char *filename, trixi_io_header *h) {
  if (!filename || !h) return fail(TRIXI_IO_EINVAL, "invalid arguments");
  Quiet quiet;
  Id f(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(TRIXI_IO_EFILE, std::string("cannot open ") + filename);
  std::string eq, mt, mf;
  const bool ok = get_i64(f.id, "ndims", &h->ndims) && get_str(f.id, "equations", &eq) &&
                  get_i64(f.id, "polydeg", &h->polydeg) && get_i64(f.id, "n_vars", &h->n_vars) &&
                  get_i64(f.id, "n_elements", &h->n_elements) && get_str(f.id, "mesh_type", &mt) &&
                  get_str(f.id, "mesh_file", &mf) && get_f64(f.id, "time", &h->time) &&
                  get_f64(f.id, "dt", &h->dt) && get_i64(f.id, "timestep", &h->timestep);
  if (!ok) return fail(TRIXI_IO_EMISMATCH, std::string(filename) + " lacks a solution/restart attribute");
  if (!copy_str(eq, h->equations, sizeof(h->equations)) || !copy_str(mt, h->mesh_type, sizeof(h->mesh_type)) ||
      !copy_str(mf, h->mesh_file, sizeof(h->mesh_file)))
    return fail(TRIXI_IO_EMISMATCH, "string attribute too long");
  return TRIXI_IO_OK;
}

int trixi_io_read_variables(const char *filename, int64_t n_vars, int64_t n_nodes, int64_t n_elements,
                            const char *const *varnames, double *u) {
  if (!filename || n_vars < 1 || n_nodes < 1 || n_elements < 1 || !varnames || !u)
    return fail(TRIXI_IO_EINVAL, "invalid arguments");
  Quiet quiet;
  Id f(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(TRIXI_IO_EFILE, std::string("cannot open ") + filename);
  const size_t count = (size_t)n_nodes * (size_t)n_elements;
  std::vector<double> column(count);
  for (int64_t v = 0; v < n_vars; ++v) {
    const std::string name = "variables_" + std::to_string(v + 1);
    Id ds(H5Dopen2(f.id, name.c_str(), H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(TRIXI_IO_EMISMATCH, name + " missing");
    std::string vname;
    if (!get_str(ds.id, "name", &vname)) return fail(TRIXI_IO_EMISMATCH, name + " has no name attribute");
    if (!varnames[v] || vname != varnames[v])
      return fail(TRIXI_IO_EMISMATCH, "mismatch: " + name + " should be '" + (varnames[v] ? varnames[v] : "") +
                                          "', but found '" + vname + "'");
    if (!read_dataset(f.id, name.c_str(), H5T_NATIVE_DOUBLE, (hssize_t)count, column.data()))
      return fail(TRIXI_IO_EMISMATCH, name + " has the wrong size");
    for (size_t k = 0; k < count; ++k) u[(size_t)v + (size_t)n_vars * k] = column[k];
  }
  return TRIXI_IO_OK;
}

int trixi_io_write_tree(const char *filename, const trixi_io_tree_header *h, const int64_t *parent_ids,
                        const int64_t *child_ids, const int64_t *neighbor_ids, const int64_t *levels,
                        const double *coordinates) {
  if (!filename || !h || (h->ndims != 2 && h->ndims != 3) || h->n_cells < 1 || !parent_ids || !child_ids ||
      !neighbor_ids || !levels || !coordinates)
    return fail(TRIXI_IO_EINVAL, "invalid arguments");
  Quiet quiet;
  Id f(H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(TRIXI_IO_EFILE, std::string("cannot create ") + filename);
  const hsize_t d = (hsize_t)h->ndims, n = (hsize_t)h->n_cells;
  bool ok = put_str(f.id, "mesh_type", "TreeMesh") && put_i64(f.id, "ndims", h->ndims) &&
            put_i64(f.id, "n_cells", h->n_cells) && put_i64(f.id, "n_leaf_cells", h->n_leaf_cells) &&
            put_i64(f.id, "minimum_level", h->minimum_level) && put_i64(f.id, "maximum_level", h->maximum_level) &&
            put_attr(f.id, "center_level_0", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 1, &d, h->center_level_0) &&
            put_f64(f.id, "length_level_0", h->length_level_0) &&
            put_attr(f.id, "periodicity", H5T_STD_B8LE, H5T_NATIVE_B8, 1, &d, h->periodicity);
  const hsize_t dims_child[2] = {n, (hsize_t)1 << d}, dims_nbr[2] = {n, 2 * d}, dims_coord[2] = {n, d};
  ok = ok && put_dataset(f.id, "parent_ids", H5T_STD_I64LE, H5T_NATIVE_INT64, 1, &n, parent_ids) &&
       put_dataset(f.id, "child_ids", H5T_STD_I64LE, H5T_NATIVE_INT64, 2, dims_child, child_ids) &&
       put_dataset(f.id, "neighbor_ids", H5T_STD_I64LE, H5T_NATIVE_INT64, 2, dims_nbr, neighbor_ids) &&
       put_dataset(f.id, "levels", H5T_STD_I64LE, H5T_NATIVE_INT64, 1, &n, levels) &&
       put_dataset(f.id, "coordinates", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, dims_coord, coordinates);
  if (!ok) return fail(TRIXI_IO_EFILE, std::string("writing ") + filename + " failed");
  return TRIXI_IO_OK;
}

int trixi_io_read_tree_header(const char *filename, trixi_io_tree_header *h) {
  if (!filename || !h) return fail(TRIXI_IO_EINVAL, "invalid arguments");
  Quiet quiet;
  Id f(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(TRIXI_IO_EFILE, std::string("cannot open ") + filename);
  std::string mt;
  if (!get_str(f.id, "mesh_type", &mt) || mt != "TreeMesh")
    return fail(TRIXI_IO_EMISMATCH, std::string(filename) + " is not a TreeMesh mesh file");
  std::memset(h, 0, sizeof(*h));
  if (!get_i64(f.id, "ndims", &h->ndims) || (h->ndims != 2 && h->ndims != 3))
    return fail(TRIXI_IO_EMISMATCH, "ndims missing or unsupported");
  const bool ok = get_i64(f.id, "n_cells", &h->n_cells) && get_i64(f.id, "n_leaf_cells", &h->n_leaf_cells) &&
                  get_i64(f.id, "minimum_level", &h->minimum_level) &&
                  get_i64(f.id, "maximum_level", &h->maximum_level) &&
                  get_f64_vec(f.id, "center_level_0", h->center_level_0, h->ndims) &&
                  get_f64(f.id, "length_level_0", &h->length_level_0) &&
                  get_bool_vec(f.id, "periodicity", h->periodicity, h->ndims);
  if (!ok) return fail(TRIXI_IO_EMISMATCH, std::string(filename) + " lacks a mesh attribute");
  return TRIXI_IO_OK;
}

int trixi_io_read_tree(const char *filename, int64_t n_cells, int64_t ndims, int64_t *parent_ids,
                       int64_t *child_ids, int64_t *neighbor_ids, int64_t *levels, double *coordinates) {
  if (!filename || n_cells < 1 || (ndims != 2 && ndims != 3) || !parent_ids || !child_ids || !neighbor_ids ||
      !levels || !coordinates)
    return fail(TRIXI_IO_EINVAL, "invalid arguments");
  Quiet quiet;
  Id f(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) return fail(TRIXI_IO_EFILE, std::string("cannot open ") + filename);
  const hssize_t n = n_cells;
  const bool ok = read_dataset(f.id, "parent_ids", H5T_NATIVE_INT64, n, parent_ids) &&
                  read_dataset(f.id, "child_ids", H5T_NATIVE_INT64, n << ndims, child_ids) &&
                  read_dataset(f.id, "neighbor_ids", H5T_NATIVE_INT64, n * 2 * ndims, neighbor_ids) &&
                  read_dataset(f.id, "levels", H5T_NATIVE_INT64, n, levels) &&
                  read_dataset(f.id, "coordinates", H5T_NATIVE_DOUBLE, n * ndims, coordinates);
  if (!ok) return fail(TRIXI_IO_EMISMATCH, std::string(filename) + ": a tree dataset is missing or has the wrong size");
  return TRIXI_IO_OK;
}

}  // extern "C"
