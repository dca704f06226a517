/* HDF5 block output and restart input of a block's fields, in the layout of
 * the files Cello's OutputData writes (src/Cello/io_OutputData.cpp:113-250)
 * and IoEnzoReader reads back (src/Enzo/io/control_restart.cpp:559-720):
 *
 *   <dir>/<file>                 one HDF5 file (libhdf5 1.10 C API)
 *     attributes  rank cycle time dt          IoSimulation (io_IoSimulation.cpp)
 *                 lower upper max_level       IoHierarchy  (io_IoHierarchy.cpp)
 *                 version                     cello::io::write_version_metadata
 *     /B<bits>    one group per block, named by Block::name
 *       attributes  num_field_data index lower upper cycle time dt array
 *                   index_order count_order           IoBlock (io_IoBlock.cpp:12-90)
 *                   enzo_dt enzo_GridLeftEdge enzo_GridDimension
 *                   enzo_GridStartIndex enzo_GridEndIndex enzo_CellWidth
 *                   enzo_redshift                     IoEnzoBlock (IoEnzoBlock.cpp)
 *       field_<name>  native double dataset (mz, my, mx), ghost zones
 *                     included (IoFieldData::field_array, io_IoFieldData.cpp:
 *                     60-130: include_ghosts_ is always true)
 *   <dir>/<list>.block_list      "<block name> <file>" per block
 *   <dir>/<list>.file_list       "<file>" per file
 *
 * Every attribute is a one-dimensional array (scalars have one element),
 * as FileHdf5::write_meta_ creates them (disk_FileHdf5.cpp:807-880).
 *
 * Field data may live in device memory: it then moves in slabs of whole
 * planes (32 MiB at most) through two pinned host buffers, the copy of one
 * slab overlapping the file write of the previous one, on the caller's
 * stream (after the work queued on it).  The pinned buffers are allocated
 * once per process and shared by its calls, which take turns.
 *
 * Link with -lenzoe_io (HDF5 from /opt/conda/lib).  Status codes are those
 * of enzoe_vlct.h; ENZOE_ERR_IO reports an HDF5 or file-system failure.
 */
#ifndef ENZOE_OUTPUT_H
#define ENZOE_OUTPUT_H

#include "enzoe_vlct.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ENZOE_ERR_IO 6 /* HDF5 / file-system failure                      */

/* Cello's Index of a block (src/Cello/mesh_Index.hpp:16-40): its
 * refinement level (0 = root level, at most 20) and its position on that
 * level's grid of blocks, global[a] in [0, root_blocks[a] << level)
 * (Block::index_global, src/Cello/mesh_Block.cpp:902-932).                   */
typedef struct {
  int level;
  int global[3];
} enzoe_block_index;

/* Block::name(Index) (src/Cello/mesh_Block.cpp:798-821) with
 * Index::bit_string (src/Cello/mesh_Index.cpp:544-575): "B", then per axis
 * the root-array bits (as many as root_blocks[a]-1 needs, most significant
 * first), ":" and the child bits of levels 1..level when level > 0, axes
 * separated by "_".  Writes at most out_len bytes including the NUL.        */
int enzoe_block_name(const enzoe_block_index *index, const int root_blocks[3],
                     int rank, char *out, int out_len);

/* The three packed 32-bit words of the Index (the "index" attribute):
 * NodeBits {level:2, array:10, tree:20} per axis (mesh_Index.hpp:16-40)
 * filled by Index::set_array, set_level and set_child (mesh_Index.cpp:
 * 340-411).                                                                  */
int enzoe_block_index_words(const enzoe_block_index *index,
                            const int root_blocks[3], int words[3]);

/* Output::expand_name_ (src/Cello/io_Output.cpp:125-207): replaces the
 * printf conversions of `format`, right to left, with the variables named
 * in args ("cycle", "count", "proc", "flipflop": int; "time": double).
 * The number of '%' must equal nargs.  A conversion that does not match
 * its variable's type is rejected (ENZOE_ERR_CONFIG).                        */
int enzoe_output_expand_name(const char *format, const char *const *args,
                             int nargs, int cycle, double time, int count,
                             int proc, char *out, int out_len);

/* One field of one block: `data` holds the (mz, my, mx) array with its
 * ghost zones, x fastest; face-centred fields are one longer along their
 * axis.  Source of a write, destination of a read.                          */
typedef struct {
  const char *name; /* field name; the dataset is "field_<name>"         */
  double *data;
  int mx, my, mz;
} enzoe_output_field;

/* The domain and the simulation state written once per file.             */
typedef struct {
  int rank;               /* dimensionality, 1..3                          */
  double lower[3];        /* Domain:lower                                  */
  double upper[3];        /* Domain:upper                                  */
  int root_blocks[3];     /* Mesh:root_blocks                              */
  int max_level;          /* Adapt:max_level                               */
  int cycle;              /* Simulation state (IoSimulation)               */
  double time;
  double dt;
} enzoe_output_hierarchy;

/* One block to write.                                                      */
typedef struct {
  enzoe_block_index index;
  int cycle;              /* Block state (IoBlock)                         */
  double time;
  double dt;
  int n[3];               /* active cells per axis                         */
  int ghost[3];           /* ghost depth per axis                          */
  const enzoe_output_field *fields;
  int nfields;
  int on_device;          /* nonzero: field data in device memory          */
} enzoe_output_block;

typedef struct enzoe_output_s *enzoe_output_t;

/* OutputData::open + write_hierarchy (io_OutputData.cpp:68-122): expands
 * the file and directory names (dir_format NULL or "" = the working
 * directory "."), creates the directory and the HDF5 file, writes the
 * file attributes and opens the block and file lists.                      */
int enzoe_output_open(const char *file_format, const char *const *file_args,
                      int n_file_args, const char *dir_format,
                      const char *const *dir_args, int n_dir_args, int count,
                      int proc, const enzoe_output_hierarchy *h,
                      enzoe_output_t *out);

/* OutputData::write_block + write_field_data (io_OutputData.cpp:126-250):
 * the block's group, its attributes and one dataset per field, in the
 * order given.  Returns after the data is in the file.                     */
int enzoe_output_write_block(enzoe_output_t o, const enzoe_output_block *b,
                             enzoe_stream_t stream);

/* Name of the HDF5 file written (valid until enzoe_output_close).          */
const char *enzoe_output_file_path(enzoe_output_t o);

/* OutputData::close: closes the HDF5 file and the lists, frees the handle. */
int enzoe_output_close(enzoe_output_t o);

/* IoEnzoReader::file_read_hierarchy_ (control_restart.cpp:590-622): the
 * simulation state of a file written above.                                */
int enzoe_input_read_state(const char *path, int *cycle, double *time,
                           double *dt);

/* IoEnzoReader::file_read_block_fields_ (control_restart.cpp:667-720): the
 * datasets field_<name> of group /<block_name> into `fields` (host or
 * device memory); each dataset's extents must equal the field's.
 * cycle/time/dt (may be NULL) receive the block's state attributes.        */
int enzoe_input_read_block(const char *path, const char *block_name,
                           const enzoe_output_field *fields, int nfields,
                           int on_device, enzoe_stream_t stream, int *cycle,
                           double *time, double *dt);

#ifdef __cplusplus
}
#endif

#endif
