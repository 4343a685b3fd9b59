// acc_driver.cpp -- example C++ host program over the adapter (vacc_host.hpp): the per-step
// loop of calculateAcceleration (vlasovmover.cpp:410-475) for a set of spatial cells of one
// population: per cell the FP64 transform and intersections, then the three remaps (batched,
// or one cell per launch), then the block adjust: single-cell (adjust = 1), or with spatial
// neighbours (adjust = 2: the cells form a ring, cell i neighbouring i-1 and i+1), or with
// spatial neighbours followed by the spatial translation along dimension step % 3 of the ring
// as one periodic pencil (adjust = 3; cell width 1e7 m, dt 1.5 s).  At the end the cells'
// blocks go through a VLSV file (<output>.vlsv, written and read back by the adapter).
//
//   vacc_acc_driver <input> <output>
//
// input (little endian): "VACC" u32 version(1) u32 ncells i32 method i32 steps i32 per_cell i32 adjust
//   mesh u32[3] len, f64[3] vmin, f64[3] vmax; species f64 mass, charge, min_value, i32 add_width
//   per cell: f64 B[3], V[3], dt; u32 n; u32 gids[n]; f32 f[64 n]
// output: per step per cell f64 intersections[12]; then per cell u32 n, u32 gids[n], f32 f[64 n],
//   f64 RHOLOSSADJUST; then i32 OR of the status codes.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vacc_host.hpp"

template <class T>
static bool rd(FILE* fp, T* p, size_t n = 1) {
   return fread(p, sizeof(T), n, fp) == n;
}
template <class T>
static void wr(FILE* fp, const T* p, size_t n = 1) {
   fwrite(p, sizeof(T), n, fp);
}

int main(int argc, char** argv) {
   if (argc != 3) {
      fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
      return 2;
   }
   FILE* in = fopen(argv[1], "rb");
   if (!in) return 2;
   char magic[4];
   uint32_t version, ncells;
   int32_t method, steps, per_cell, do_adjust;
   vacc_mesh mesh{};
   vacc_species sp{};
   if (!rd(in, magic, 4) || memcmp(magic, "VACC", 4) || !rd(in, &version) || version != 1 || !rd(in, &ncells) ||
       !rd(in, &method) || !rd(in, &steps) || !rd(in, &per_cell) || !rd(in, &do_adjust) ||
       !rd(in, mesh.grid_length, 3) || !rd(in, mesh.min_limits, 3) || !rd(in, mesh.max_limits, 3) ||
       !rd(in, &sp.mass) || !rd(in, &sp.charge) || !rd(in, &sp.sparse_min_value) || !rd(in, &sp.sparse_add_width_v)) {
      fprintf(stderr, "bad input header\n");
      return 2;
   }
   std::vector<vacc_host::SpatialCell> cells(ncells);
   std::vector<double> dt(ncells);
   for (auto& c : cells) {
      const size_t i = &c - cells.data();
      c.populations.resize(1);
      uint32_t n = 0;
      if (!rd(in, c.B, 3) || !rd(in, c.V, 3) || !rd(in, &dt[i]) || !rd(in, &n)) return 2;
      c.populations[0].gids.resize(n);
      c.populations[0].f.resize((size_t)n * VACC_WID3);
      if (!rd(in, c.populations[0].gids.data(), n) || !rd(in, c.populations[0].f.data(), (size_t)n * VACC_WID3))
         return 2;
   }
   fclose(in);

   vacc_options opts{};
   opts.method = method;
   opts.storage = VACC_STORE_F32;
   opts.hall_minimum_rhoq = 1.60217653e-19;
   opts.max_rotation_deg = 25.0;
   vacc_host::AccelerationEngine engine({mesh}, {sp}, opts);
   std::vector<vacc_host::SpatialCell*> ptrs;
   for (auto& c : cells) ptrs.push_back(&c);

   FILE* out = fopen(argv[2], "wb");
   if (!out) return 2;
   int status_or = 0;
   for (int s = 0; s < steps; ++s) {
      const uint32_t map_order = (uint32_t)(s % 3);
      for (uint32_t i = 0; i < ncells; ++i) {
         const int rc = engine.compute_cell_intersections(ptrs[i], 0, map_order, dt[i]);
         if (rc < 0) {
            fprintf(stderr, "intersections: %s\n", engine.last_error().c_str());
            return 1;
         }
         wr(out, cells[i].populations[0].intersection, 12);
      }
      int rc = 0;
      if (per_cell) {
         for (auto* c : ptrs) {
            rc = engine.accelerate_cell(c, 0, map_order);
            if (rc < 0) break;
            status_or |= rc;
         }
      } else {
         rc = engine.accelerate_cells(ptrs, 0, map_order);
         if (rc > 0) status_or |= rc;
      }
      if (rc >= 0 && do_adjust == 1) rc = engine.adjust_blocks(ptrs, 0);
      if (rc >= 0 && (do_adjust == 2 || do_adjust == 3)) {
         std::vector<std::vector<uint32_t>> ring(ncells);
         for (uint32_t i = 0; i < ncells; ++i) ring[i] = {(i + ncells - 1) % ncells, (i + 1) % ncells};
         rc = engine.adjust_blocks(ptrs, 0, ring);
      }
      if (rc >= 0 && do_adjust == 3) {
         // the ring as one periodic pencil: stencil cells n-2, n-1 | 0 .. n-1 | 0, 1
         vacc_host::PencilSet ps;
         const uint32_t L = ncells + 2 * VACC_TRANS_STENCIL;
         ps.lengthOfPencils = {L};
         ps.idsStart = {0};
         for (uint32_t q = 0; q < L; ++q) {
            ps.ids.push_back((q + 2 * ncells - VACC_TRANS_STENCIL) % ncells);
            ps.sourceDZ.push_back(1.0e7f);
            ps.targetRatios.push_back((q >= VACC_TRANS_STENCIL - 1 && q <= L - VACC_TRANS_STENCIL) ? 1.0f : 0.0f);
         }
         rc = engine.trans_map_1d(ptrs, ps, map_order, 1.5, 0);
      }
      if (rc < 0) {
         fprintf(stderr, "acceleration: %s\n", engine.last_error().c_str());
         return 1;
      }
   }
   // through a VLSV restart file and back
   {
      const std::string path = std::string(argv[2]) + ".vlsv";
      std::vector<uint64_t> ids(ncells);
      for (uint32_t i = 0; i < ncells; ++i) ids[i] = 1000 + i;
      vacc_vlsv_writer* w = nullptr;
      if (vacc_vlsv_create(path.c_str(), &w) || engine.write_velocity_distribution(0, w, ptrs, ids, "proton") < 0 ||
          vacc_vlsv_close(w)) {
         fprintf(stderr, "vlsv write: %s\n", engine.last_error().c_str());
         return 1;
      }
      for (auto& c : cells) {
         c.populations[0].gids.clear();
         c.populations[0].f.clear();
      }
      vacc_vlsv_reader* r = nullptr;
      std::vector<uint64_t> got;
      if (vacc_vlsv_open(path.c_str(), &r) || engine.read_block_data(r, "SpatialGrid", 0, ptrs, got, 0, "proton") < 0) {
         fprintf(stderr, "vlsv read: %s\n", engine.last_error().c_str());
         return 1;
      }
      vacc_vlsv_reader_close(r);
      if (got != ids) {
         fprintf(stderr, "vlsv read: cell ids differ\n");
         return 1;
      }
   }
   for (auto& c : cells) {
      const auto& p = c.populations[0];
      const uint32_t n = (uint32_t)p.gids.size();
      wr(out, &n);
      wr(out, p.gids.data(), n);
      wr(out, p.f.data(), (size_t)n * VACC_WID3);
      wr(out, &p.RHOLOSSADJUST);
   }
   wr(out, &status_or);
   fclose(out);
   return 0;
}
