/* trixi_hip_io.h -- solution, restart and mesh files of Trixi.jl's TreeMesh DG
 * solvers (HDF5), host-only companion library of libtrixi_hip.
 *
 * The files carry the same attributes and datasets, with the same names, types
 * and shapes, as the reference writes through HDF5.jl:
 *   save_solution_file  src/callbacks_step/save_solution_dg.jl:8-90
 *   save_restart_file   src/callbacks_step/save_restart_dg.jl:8-40
 *   load_restart_file   src/callbacks_step/save_restart_dg.jl:42-80
 *   save_mesh_file      src/meshes/mesh_io.jl:15-51 (TreeMesh)
 *   load_mesh!          src/meshes/mesh_io.jl:603-628 (TreeMesh)
 * Julia arrays are column-major; as HDF5.jl does, a [d1, d2] array is stored as
 * an HDF5 dataset of dims (d2, d1) with the same bytes. Integers are int64,
 * reals float64, strings fixed-length UTF-8, Bools 8-bit bitfields.
 *
 * Every function returns 0 on success, a negative code on failure; the message
 * of the last failure of the calling thread is trixi_io_last_error(). */
#ifndef TRIXI_HIP_IO_H
#define TRIXI_HIP_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRIXI_IO_OK 0
#define TRIXI_IO_EINVAL -1   /* bad argument */
#define TRIXI_IO_EFILE -2    /* HDF5 could not create, open, read or write */
#define TRIXI_IO_EMISMATCH -3 /* the file does not match what the caller expects */

/* file-level attributes of solution and restart files */
typedef struct {
  int64_t ndims, polydeg, n_vars, n_elements, timestep;
  double time, dt;
  char equations[64]; /* get_name(equations), e.g. "CompressibleEulerEquations3D" */
  char mesh_type[32]; /* get_name(mesh): "TreeMesh" */
  char mesh_file[256]; /* file name (no directory) of the mesh file */
} trixi_io_header;

/* file-level attributes of a TreeMesh mesh file */
typedef struct {
  int64_t ndims, n_cells, n_leaf_cells, minimum_level, maximum_level;
  double center_level_0[3], length_level_0;
  uint8_t periodicity[3];
} trixi_io_tree_header;

const char *trixi_io_last_error(void);

/* Solution or restart file: the attributes of h, then for v = 1..h->n_vars the
 * dataset "variables_v" = u[v, :, :] flattened (n_nodes * n_elements values of
 * u [n_vars, n_nodes, n_elements], column-major) with attribute "name" =
 * varnames[v-1]; then "element_variables_k" (k = 1..n_element_variables) =
 * element_values[k-1][:] (n_elements values each) with attribute "name". */
int trixi_io_write_solution(const char *filename, const trixi_io_header *h, int64_t n_nodes, const double *u,
                            const char *const *varnames, int n_element_variables, const double *element_values,
                            const char *const *element_names);

/* The file-level attributes of a solution or restart file. */
int trixi_io_read_header(const char *filename, trixi_io_header *h);

/* load_restart_file: u [n_vars, n_nodes, n_elements] from "variables_v", after
 * checking that every dataset holds n_nodes * n_elements values and is named
 * varnames[v-1] (TRIXI_IO_EMISMATCH otherwise). */
int trixi_io_read_variables(const char *filename, int64_t n_vars, int64_t n_nodes, int64_t n_elements,
                            const char *const *varnames, double *u);

/* save_mesh_file(mesh::TreeMesh): the attributes of h and the datasets
 * parent_ids [n_cells], child_ids [2^ndims, n_cells], neighbor_ids
 * [2 ndims, n_cells], levels [n_cells], coordinates [ndims, n_cells] (1-based
 * cell ids, 0 = none; arrays given column-major as the Julia tree holds them). */
int trixi_io_write_tree(const char *filename, const trixi_io_tree_header *h, const int64_t *parent_ids,
                        const int64_t *child_ids, const int64_t *neighbor_ids, const int64_t *levels,
                        const double *coordinates);

/* The attributes of a TreeMesh mesh file (mesh_type must be "TreeMesh"). */
int trixi_io_read_tree_header(const char *filename, trixi_io_tree_header *h);

/* The datasets of a TreeMesh mesh file with n_cells cells (sizes as written). */
int trixi_io_read_tree(const char *filename, int64_t n_cells, int64_t ndims, int64_t *parent_ids,
                       int64_t *child_ids, int64_t *neighbor_ids, int64_t *levels, double *coordinates);

#ifdef __cplusplus
}
#endif

#endif /* TRIXI_HIP_IO_H */
