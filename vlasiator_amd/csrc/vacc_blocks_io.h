// vacc_blocks_io.h -- library-internal streamed transfer of a context's blocks between HBM
// and host memory, used by the VLSV block I/O.  Not part of the C ABI.
//
// Values travel in chunks of whole blocks through two pinned host buffers (and, for the
// 16-bit storage forms, a device FP32 stage that the codec kernels read or write), so that
// the copy of one chunk overlaps the host-side consumer or producer of the other.
#pragma once

#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/vacc.h"

#define VACC_INTERNAL __attribute__((visibility("hidden")))

namespace vacc {

// Receives n blocks (64 floats each, in-block index i + 4j + 16k) starting at block `first`
// of the exported list.  Returns 0 or a negative error code, which ends the transfer.
using BlockSink = std::function<int(const float* f, uint64_t first, uint64_t n)>;
// Fills n blocks starting at block `first` of the list being imported.
using BlockSource = std::function<int(float* f, uint64_t first, uint64_t n)>;

// Sets the message returned by vacc_last_error and returns code.
VACC_INTERNAL int set_error(int code, const std::string& msg);

// The context's mesh and its cell size per axis (velocity_mesh_parameters.cpp:153-163).
VACC_INTERNAL void ctx_mesh(const vacc_ctx* c, vacc_mesh* mesh, double cell[3]);

// Block counts per cell and block GIDs (ascending per cell) of the current state; leaves the
// state compacted so that export_values streams the blocks in the same order.
VACC_INTERNAL int export_lists(vacc_ctx* c, std::vector<uint32_t>& counts, std::vector<uint32_t>& gids);
// Streams the values of the blocks listed by the preceding export_lists, as FP32.
VACC_INTERNAL int export_values(vacc_ctx* c, const BlockSink& sink);

// Replaces the context's cells by ncells cells with blocks [offsets[c], offsets[c+1]) of
// gids; the values come from `source` (same contract and checks as vacc_upload_cells).
VACC_INTERNAL int import_blocks(vacc_ctx* c, uint32_t ncells, const uint64_t* offsets, const uint32_t* gids,
                                const BlockSource& source);

}  // namespace vacc
