// vacc_host.hpp -- C++ host adapter over the C ABI of libvacc (include/vacc.h).
//
// Mirrors the part of the reference's host interface that the acceleration path uses, so
// that a caller written against it can switch to the GPU library:
//   Population          <- spatial_cells/spatial_cell_cpu.hpp:66-92 (velocity mesh GIDs, block
//                          data, intersection_*, subcycleDt, velocityBlockMinValue, RHOLOSSADJUST)
//   SpatialCell         <- the fields the transform reads (BGB*VOL + PERB*VOL, dPERB*VOL/d*,
//                          V_V, RHOQ_V, E*GRADPE; cpu_acc_transform.cpp:71-115)
//   compute_cell_intersections(SpatialCell*, popID, map_order, dt)
//                       <- vlasovsolver/cpu_acc_intersections.cpp:44
//   accelerate_cell(SpatialCell*, popID, map_order)
//                       <- cpu_accelerate_cell, vlasovsolver/cpu_acc_semilag.cpp:92 (one cell
//                          per launch)
//   accelerate_cells(cells, popID, map_order)
//                       <- cpu_accelerate_cells, vlasovsolver/cpu_acc_semilag.cpp:45 (all cells
//                          in one batch: every kernel covers every cell)
//   adjust_blocks(cells, popID)
//                       <- SpatialCell::adjustSingleCellVelocityBlocks(popID, true),
//                          spatial_cells/spatial_cell_cpu.cpp:275
//   adjust_blocks(cells, popID, neighbours)
//                       <- adjustVelocityBlocks(mpiGrid, cells, true, popID), grid.cpp:775-814:
//                          SpatialCell::adjust_velocity_blocks(spatial_neighbors, popID, true),
//                          spatial_cell_cpu.cpp:179-273, with the neighbours of cells[i] given as
//                          indices into cells
//   PencilSet           <- setOfPencils (vlasovsolver/cpu_trans_pencils.hpp:34-51), cell ids
//                          given as indices into cells
//   trans_map_1d(cells, pencils, dimension, dt, popID)
//                       <- trans_map_1d_amr(mpiGrid, localPropagatedCells, remoteTargetCells,
//                          nPencils, dimension, dt, popID), vlasovsolver/cpu_trans_map_amr.cpp:214
//   write_velocity_distribution(popID, writer, cells, cellIds, popName)
//                       <- writeVelocityDistributionData(popID, vlsvWriter, mpiGrid, cells, comm),
//                          iowrite.cpp:141
//   read_block_data(reader, meshName, localCellStartOffset, cells, fileCells, popID, popName)
//                       <- readBlockData(file, meshName, fileCells, localCellStartOffset,
//                          localCells, mpiGrid), ioread.cpp:488 (one population)
// Errors are reported as in the C ABI: the functions return the library's status code
// (0 ok, > 0 per-cell status bits, < 0 error) and never abort.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vacc.h"
#include "../../include/vacc_trans.h"
#include "../../include/vacc_vlsv.h"

namespace vacc_host {

struct Population {
   std::vector<uint32_t> gids;          // block GIDs in local (LID) order
   std::vector<float> f;                // 64 values per block, in-block index i + 4j + 16k
   double intersection[12] = {0};       // x {I,di,dj,dk}, y {...}, z {...}
   double subcycleDt = 0.0;
   double RHOLOSSADJUST = 0.0;
   uint32_t status = 0;                 // VACC_STATUS_* bits of the last call
};

// Pencils of one dimension, in the reference's processing order (bins, then pencils).
struct PencilSet {
   std::vector<uint32_t> lengthOfPencils;   // positions per pencil, stencil cells included
   std::vector<uint64_t> idsStart;          // first position of each pencil
   std::vector<uint32_t> ids;               // index into the cells vector per position
   std::vector<float> sourceDZ;             // cell width along the dimension per position
   std::vector<float> targetRatios;         // 0: the position is not written
};

struct SpatialCell {
   double B[3] = {0, 0, 0};             // BGB*VOL + PERB*VOL
   double dPERB[6] = {0, 0, 0, 0, 0, 0};// dPERBXVOLdy dPERBXVOLdz dPERBYVOLdx dPERBYVOLdz dPERBZVOLdx dPERBZVOLdy
   double V[3] = {0, 0, 0};             // V_V
   double RHOQ = 0.0;                   // RHOQ_V
   double EGRADPE[3] = {0, 0, 0};
   std::vector<Population> populations;
};

class AccelerationEngine {
public:
   // one velocity mesh and species per population, shared options
   AccelerationEngine(const std::vector<vacc_mesh>& meshes, const std::vector<vacc_species>& species,
                      const vacc_options& opts)
       : meshes_(meshes), species_(species), opts_(opts), ctx_(meshes.size(), nullptr) {}
   ~AccelerationEngine() {
      for (auto c : ctx_)
         if (c) vacc_destroy(c);
   }
   AccelerationEngine(const AccelerationEngine&) = delete;
   AccelerationEngine& operator=(const AccelerationEngine&) = delete;

   static std::string last_error() { return vacc_last_error(); }

   // FP64 transform + SLICE-3D intersections of one cell (host side, like the reference).
   int compute_cell_intersections(SpatialCell* sc, uint32_t popID, uint32_t map_order, double dt) {
      vacc_cell_fields fl = fields(*sc, dt);
      Population& pop = sc->populations.at(popID);
      return vacc_cell_intersections_host(&meshes_.at(popID), &species_.at(popID), &opts_, &fl, map_order,
                                          pop.intersection);
   }

   // Three 1-D remaps of one cell (one cell per launch).
   int accelerate_cell(SpatialCell* sc, uint32_t popID, uint32_t map_order) {
      std::vector<SpatialCell*> one{sc};
      return run(one, popID, map_order, false);
   }

   // Three 1-D remaps of every cell in one batch.
   int accelerate_cells(std::vector<SpatialCell*>& cells, uint32_t popID, uint32_t map_order) {
      return run(cells, popID, map_order, false);
   }

   // Single-cell sparsity adjust of every cell (doDeleteEmptyBlocks = true).
   int adjust_blocks(std::vector<SpatialCell*>& cells, uint32_t popID) {
      return run(cells, popID, 0, true);
   }

   // Sparsity adjust with spatial neighbours: blocks with content in a neighbour are kept
   // (or added) as they are.  neighbours[i] lists indices into cells.
   int adjust_blocks(std::vector<SpatialCell*>& cells, uint32_t popID,
                     const std::vector<std::vector<uint32_t>>& neighbours) {
      if (neighbours.size() != cells.size()) return VACC_E_ARG;
      return run(cells, popID, 0, true, &neighbours);
   }

   // Translation of every cell along dimension by dt over the pencils (which index cells).
   int trans_map_1d(std::vector<SpatialCell*>& cells, const PencilSet& pencils, uint32_t dimension, double dt,
                    uint32_t popID) {
      return run(cells, popID, 0, false, nullptr, &pencils, dimension, dt);
   }

   // The population's block arrays of cells (spatial cell IDs cellIds) into an open VLSV file.
   int write_velocity_distribution(uint32_t popID, vacc_vlsv_writer* writer, std::vector<SpatialCell*>& cells,
                                   const std::vector<uint64_t>& cellIds, const std::string& popName) {
      if (cellIds.size() != cells.size()) return VACC_E_ARG;
      vacc_ctx* c = nullptr;
      int rc = upload(cells, popID, &c);
      if (rc < 0) return rc;
      return vacc_vlsv_write_velocity_distribution(writer, c, popName.c_str(), cellIds.data());
   }

   // Cells [localCellStartOffset, localCellStartOffset + cells.size()) of the population from a
   // VLSV file into cells; fileCells receives their spatial cell IDs.
   int read_block_data(vacc_vlsv_reader* reader, const std::string& meshName, uint64_t localCellStartOffset,
                       std::vector<SpatialCell*>& cells, std::vector<uint64_t>& fileCells, uint32_t popID,
                       const std::string& popName) {
      vacc_ctx* c = nullptr;
      int rc = context(popID, &c);
      if (rc) return rc;
      fileCells.assign(cells.size(), 0);
      if ((rc = vacc_vlsv_read_block_data(reader, c, popName.c_str(), meshName.c_str(), localCellStartOffset,
                                          (uint32_t)cells.size(), fileCells.data())) < 0)
         return rc;
      return download(cells, popID, c, false, nullptr, nullptr);
   }

private:
   static vacc_cell_fields fields(const SpatialCell& sc, double dt) {
      vacc_cell_fields fl{};
      for (int i = 0; i < 3; ++i) {
         fl.B[i] = sc.B[i];
         fl.V[i] = sc.V[i];
         fl.EgradPe[i] = sc.EGRADPE[i];
      }
      for (int i = 0; i < 6; ++i) fl.dB[i] = sc.dPERB[i];
      fl.rhoq = sc.RHOQ;
      fl.dt = dt;
      return fl;
   }

   int context(uint32_t popID, vacc_ctx** out) {
      if (popID >= ctx_.size()) return VACC_E_ARG;
      if (!ctx_[popID]) {
         const int rc = vacc_create(&meshes_[popID], &species_[popID], &opts_, &ctx_[popID]);
         if (rc) return rc;
      }
      *out = ctx_[popID];
      return 0;
   }

   // upload the cells' populations into the population's context
   int upload(std::vector<SpatialCell*>& cells, uint32_t popID, vacc_ctx** out) {
      vacc_ctx* c = nullptr;
      int rc = context(popID, &c);
      if (rc) return rc;
      const size_t n = cells.size();
      std::vector<uint64_t> off(n + 1, 0);
      for (size_t i = 0; i < n; ++i) off[i + 1] = off[i] + cells[i]->populations.at(popID).gids.size();
      std::vector<uint32_t> gids(off[n]);
      std::vector<float> f(off[n] * VACC_WID3);
      for (size_t i = 0; i < n; ++i) {
         const Population& p = cells[i]->populations[popID];
         std::copy(p.gids.begin(), p.gids.end(), gids.begin() + off[i]);
         std::copy(p.f.begin(), p.f.end(), f.begin() + off[i] * VACC_WID3);
      }
      if ((rc = vacc_upload_cells(c, (uint32_t)n, off.data(), gids.data(), f.data())) < 0) return rc;
      *out = c;
      return 0;
   }

   // write the context's state back into the cells (status words and removed mass optional)
   int download(std::vector<SpatialCell*>& cells, uint32_t popID, vacc_ctx* c, bool adjust,
                const std::vector<uint32_t>* status, const std::vector<double>* loss) {
      const size_t n = cells.size();
      std::vector<uint32_t> counts(n);
      uint32_t nc = 0;
      int rc = vacc_get_counts(c, &nc, counts.data());
      if (rc < 0) return rc;
      if (nc != n) return VACC_E_STATE;
      uint64_t tot = 0;
      for (uint32_t v : counts) tot += v;
      std::vector<uint64_t> off(n + 1, 0);
      std::vector<uint32_t> gids(tot);
      std::vector<float> f(tot * VACC_WID3);
      if ((rc = vacc_download_cells(c, off.data(), gids.data(), f.data())) < 0) return rc;
      for (size_t i = 0; i < n; ++i) {
         Population& p = cells[i]->populations.at(popID);
         p.gids.assign(gids.begin() + off[i], gids.begin() + off[i + 1]);
         p.f.assign(f.begin() + off[i] * VACC_WID3, f.begin() + off[i + 1] * VACC_WID3);
         if (status) p.status = (*status)[i];
         if (adjust && loss) p.RHOLOSSADJUST += (*loss)[i];
      }
      return 0;
   }

   // upload the cells' populations, remap, adjust or translate, and write the new state back
   int run(std::vector<SpatialCell*>& cells, uint32_t popID, uint32_t map_order, bool adjust,
           const std::vector<std::vector<uint32_t>>* neighbours = nullptr, const PencilSet* pencils = nullptr,
           uint32_t dimension = 0, double dt = 0.0) {
      vacc_ctx* c = nullptr;
      int rc = upload(cells, popID, &c);
      if (rc < 0) return rc;
      const size_t n = cells.size();
      int st = 0;
      std::vector<double> loss(n, 0.0);
      if (pencils) {
         const PencilSet& P = *pencils;
         const size_t np = P.lengthOfPencils.size(), npos = P.ids.size();
         if (P.idsStart.size() != np || P.sourceDZ.size() != npos || P.targetRatios.size() != npos)
            return VACC_E_ARG;
         vacc_pencils vp{(uint32_t)np, P.lengthOfPencils.data(), P.idsStart.data(), P.ids.data(), P.sourceDZ.data(),
                         P.targetRatios.data()};
         if ((rc = vacc_trans_set_pencils(c, dimension, &vp)) < 0) return rc;
         st = vacc_trans_map_1d(c, dimension, dt);
      } else if (adjust && neighbours) {
         std::vector<uint32_t> noff(n + 1, 0), nbr;
         for (size_t i = 0; i < n; ++i) {
            for (uint32_t j : (*neighbours)[i]) {
               if (j >= n) return VACC_E_ARG;
               nbr.push_back(j);
            }
            noff[i + 1] = (uint32_t)nbr.size();
         }
         if (nbr.empty()) nbr.push_back(0);
         st = vacc_adjust_blocks_spatial(c, noff.data(), nbr.data(), nullptr, nullptr, loss.data());
      } else if (adjust) {
         st = vacc_adjust_blocks(c, loss.data());
      } else {
         std::vector<double> isect(12 * n);
         for (size_t i = 0; i < n; ++i) {
            const Population& p = cells[i]->populations[popID];
            std::copy(p.intersection, p.intersection + 12, isect.begin() + 12 * i);
         }
         if ((rc = vacc_set_intersections(c, isect.data())) < 0) return rc;
         st = vacc_accelerate_cells(c, map_order);
      }
      if (st < 0) return st;
      std::vector<uint32_t> status(n, 0);
      if ((rc = vacc_get_status(c, status.data())) < 0) return rc;
      if ((rc = download(cells, popID, c, adjust, &status, &loss)) < 0) return rc;
      return st;
   }

   std::vector<vacc_mesh> meshes_;
   std::vector<vacc_species> species_;
   vacc_options opts_;
   std::vector<vacc_ctx*> ctx_;
};

}  // namespace vacc_host
