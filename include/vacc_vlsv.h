/*
 * vacc_vlsv.h -- VLSV files for libvacc: the array container, and the velocity block
 * arrays of one population written from and read into a context (restart and output
 * files of fmihpc/vlasiator).  Same return convention as vacc.h (0, or VACC_E_*).
 *
 * File layout (little-endian): bytes 0-7 an endianness word (byte 0 = 0: little-endian),
 * bytes 8-15 the uint64 byte offset of the XML footer, then the arrays back to back, each
 * arraysize * vectorsize * datasize bytes, then the footer: a <VLSV> root holding one tag
 * per array whose attributes are arraysize, vectorsize, datatype (int, uint or float),
 * datasize and the writer's own (name, mesh, ...), and whose text is the array's byte
 * offset.  An array is found by its tag and the attributes the reader asks for.  Files
 * whose first word is itself the footer offset (the layout doc/vlsv2/documentation.lyx
 * describes) are read as well.
 *
 * Reference interfaces replaced (file:line in fmihpc/vlasiator):
 *   vacc_vlsv_write_velocity_distribution
 *        <- writeVelocityDistributionData(popID, vlsv::Writer&, mpiGrid, cells, comm)
 *           iowrite.cpp:141 (the arrays CELLSWITHBLOCKS, BLOCKSPERCELL, MESH_BBOX,
 *           MESH_NODE_CRDS_X/Y/Z, BLOCKIDS and BLOCKVARIABLE, :165-309)
 *   vacc_vlsv_read_block_data
 *        <- readBlockData(vlsv::ParallelReader&, meshName, fileCells, localCellStartOffset,
 *           localCells, mpiGrid) ioread.cpp:488 and _readBlockData ioread.cpp:371
 *           (velocity-mesh growth by GID renumbering, ioread.cpp:520-631; f read as float,
 *           double, int or uint of 4 or 8 bytes, ioread.cpp:667-700)
 *   vacc_vlsv_create / vacc_vlsv_write_array / vacc_vlsv_close
 *        <- vlsv::Writer::open / writeArray / close of the fmihpc/vlsv library as called at
 *           iowrite.cpp:165-309 (the library is not part of the reference tree)
 *   vacc_vlsv_open / vacc_vlsv_array_info / vacc_vlsv_read_array / vacc_vlsv_reader_close
 *        <- vlsv::ParallelReader::open / getArrayInfo / readArray / close as called at
 *           ioread.cpp:401-440
 */
#ifndef VACC_VLSV_H
#define VACC_VLSV_H

#include <stdint.h>

#include "vacc.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { VACC_VLSV_INT = 0, VACC_VLSV_UINT = 1, VACC_VLSV_FLOAT = 2 };

typedef struct vacc_vlsv_writer vacc_vlsv_writer;
typedef struct vacc_vlsv_reader vacc_vlsv_reader;

/* Creates (truncates) the file and writes the header. */
int vacc_vlsv_create(const char* path, vacc_vlsv_writer** out);
/* Appends one array: arraysize vectors of vectorsize components of datasize bytes each
 * (datatype VACC_VLSV_*; float: 4 or 8 bytes, int/uint: 1, 2, 4 or 8).  attr holds nattr
 * key/value string pairs (key0, value0, key1, ...); they may not redefine arraysize,
 * vectorsize, datatype or datasize, and may not contain the characters < > & ". */
int vacc_vlsv_write_array(vacc_vlsv_writer* w, const char* tag, uint32_t nattr, const char* const* attr,
                          uint64_t arraysize, uint64_t vectorsize, int32_t datatype, uint64_t datasize,
                          const void* data);
/* Writes the block arrays of the context's population pop_name, its cells taken in the
 * context's order with spatial cell IDs cell_ids[ncells] (NULL: 1, 2, ..., ncells).
 * Blocks of a cell are written in ascending GID order and f as FP32 (the 16-bit storage
 * forms are decoded).  The context's state is unchanged. */
int vacc_vlsv_write_velocity_distribution(vacc_vlsv_writer* w, vacc_ctx* ctx, const char* pop_name,
                                          const uint64_t* cell_ids);
/* Writes the footer and its offset, closes the file and frees w (also on error). */
int vacc_vlsv_close(vacc_vlsv_writer* w);

/* Opens a file and parses its footer. */
int vacc_vlsv_open(const char* path, vacc_vlsv_reader** out);
/* Size and type of the first array with this tag and these attributes. */
int vacc_vlsv_array_info(vacc_vlsv_reader* r, const char* tag, uint32_t nattr, const char* const* attr,
                         uint64_t* arraysize, uint64_t* vectorsize, int32_t* datatype, uint64_t* datasize);
/* Copies vectors [begin, begin + amount) of that array, as stored, to out. */
int vacc_vlsv_read_array(vacc_vlsv_reader* r, const char* tag, uint32_t nattr, const char* const* attr,
                         uint64_t begin, uint64_t amount, void* out);
/* Replaces the context's cells by cells [cell_offset, cell_offset + ncells) of population
 * pop_name on spatial mesh mesh_name ("SpatialGrid" in vlasiator files) and returns their
 * spatial cell IDs in cell_ids_out[ncells] (may be NULL).  A file velocity mesh smaller than
 * the context's is placed into it by GID renumbering when its node coordinates fall on the
 * context's cell boundaries at a block offset; a larger one is an error. */
int vacc_vlsv_read_block_data(vacc_vlsv_reader* r, vacc_ctx* ctx, const char* pop_name, const char* mesh_name,
                              uint64_t cell_offset, uint32_t ncells, uint64_t* cell_ids_out);
int vacc_vlsv_reader_close(vacc_vlsv_reader* r);

#ifdef __cplusplus
}
#endif

#endif
