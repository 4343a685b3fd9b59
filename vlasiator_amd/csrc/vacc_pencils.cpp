// vacc_pencils.cpp -- pencils of a spatial grid without refinement (include/vacc_trans.h).
//
// Host C++ only.  For refinement level 0 the reference's pencil construction
// (cpu_trans_pencils.cpp) reduces to: seed cells (getSeedIds :920-981), a walk in +dim
// from each seed until a seed, an unpropagated cell or the grid edge
// (buildPencilsWithNeighbors :719-905), VACC_TRANS_STENCIL stencil positions at each end
// holding the neighbours at distance 1, 2 (replaced by the nearest good cell when absent or
// inactive), cell widths and target ratios (computeSpatialSourceCellsForPencil :447-649).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vacc_trans.h"
#include "vacc_blocks_io.h"

using vacc::set_error;

namespace {

struct Grid {
   uint32_t n[3];
   bool periodic[3];
   const uint8_t* kind;
   uint32_t index(const uint32_t i[3]) const { return i[0] + n[0] * (i[1] + n[1] * i[2]); }
   void coords(uint32_t c, uint32_t i[3]) const {
      i[0] = c % n[0];
      i[1] = (c / n[0]) % n[1];
      i[2] = c / (n[0] * n[1]);
   }
   // do_translate_cell (cpu_trans_pencils.cpp:26-33)
   bool propagated(uint32_t c) const { return kind[c] != VACC_CELL_INACTIVE; }
   // neighbour of c at signed distance s along d, or INVALID (beyond a non-periodic edge)
   uint32_t neighbour(uint32_t c, int d, int s) const {
      uint32_t i[3];
      coords(c, i);
      const int64_t x = (int64_t)i[d] + s;
      if (x < 0 || x >= (int64_t)n[d]) {
         if (!periodic[d]) return VACC_INVALID_ID;
         i[d] = (uint32_t)(((x % (int64_t)n[d]) + n[d]) % n[d]);
      } else {
         i[d] = (uint32_t)x;
      }
      return index(i);
   }
};

}  // namespace

extern "C" int vacc_build_pencils_uniform(const uint32_t dims[3], const int32_t periodic[3], const uint8_t* cell_kind,
                                          const double dx[3], uint32_t dim, uint32_t* npencils, uint64_t* npositions,
                                          uint32_t* length, uint64_t* start, uint32_t* cell, float* dz, float* ratio) {
   if (!dims || !periodic || !cell_kind || !dx || !npencils || !npositions) return set_error(VACC_E_ARG, "null argument");
   if (dim > 2) return set_error(VACC_E_ARG, "dim must be 0, 1 or 2");
   Grid g;
   uint64_t ncells = 1;
   for (int d = 0; d < 3; ++d) {
      if (dims[d] == 0) return set_error(VACC_E_ARG, "empty grid");
      g.n[d] = dims[d];
      g.periodic[d] = periodic[d] != 0;
      ncells *= dims[d];
   }
   if (ncells >= VACC_INVALID_ID) return set_error(VACC_E_ARG, "grid too large");
   g.kind = cell_kind;
   const int W = VACC_TRANS_STENCIL;
   const int d = (int)dim;
   // seeds: the -d neighbour is across a periodic boundary, inactive, or absent
   std::vector<uint8_t> seed(ncells, 0);
   for (uint32_t c = 0; c < ncells; ++c) {
      if (!g.propagated(c)) continue;
      uint32_t i[3];
      g.coords(c, i);
      const uint32_t nb = g.neighbour(c, d, -1);
      if (nb == VACC_INVALID_ID) {
         seed[c] = 1;
      } else {
         uint32_t j[3];
         g.coords(nb, j);
         if (i[d] < j[d] || !g.propagated(nb)) seed[c] = 1;
      }
   }
   // pencils from the seeds, in cell order
   std::vector<std::vector<uint32_t>> pencils;
   for (uint32_t c = 0; c < ncells; ++c) {
      if (!seed[c]) continue;
      std::vector<uint32_t> ids(W, VACC_INVALID_ID);
      ids.push_back(c);
      uint32_t id = c;
      for (;;) {
         const uint32_t nx = g.neighbour(id, d, +1);
         if (nx == VACC_INVALID_ID || seed[nx] || !g.propagated(nx)) break;
         ids.push_back(nx);
         id = nx;
      }
      for (int s = 0; s < W; ++s) ids.push_back(VACC_INVALID_ID);
      const int L = (int)ids.size();
      // stencil positions: neighbours at distance 1..W beyond each end
      for (int s = 1; s <= W; ++s) {
         const uint32_t f = g.neighbour(ids[W], d, -s);
         if (f != VACC_INVALID_ID && ids[W - s + 1] != f) ids[W - s] = f;
         const uint32_t b = g.neighbour(ids[L - W - 1], d, +s);
         if (b != VACC_INVALID_ID && ids[L - W + s - 2] != b) ids[L - W + s - 1] = b;
      }
      // replace absent or inactive stencil cells by the nearest good cell
      uint32_t last = ids[W];
      for (int q = W - 1; q >= 0; --q) {
         if (ids[q] == VACC_INVALID_ID || !g.propagated(ids[q])) ids[q] = last;
         else last = ids[q];
      }
      last = ids[L - W - 1];
      for (int q = L - W; q < L; ++q) {
         if (ids[q] == VACC_INVALID_ID || !g.propagated(ids[q])) ids[q] = last;
         else last = ids[q];
      }
      pencils.push_back(std::move(ids));
   }
   uint64_t total = 0;
   for (const auto& p : pencils) total += p.size();
   *npencils = (uint32_t)pencils.size();
   *npositions = total;
   if (!length) return 0;
   if (!start || !cell || !dz || !ratio) return set_error(VACC_E_ARG, "null output array");
   uint64_t pos = 0;
   for (size_t p = 0; p < pencils.size(); ++p) {
      const auto& ids = pencils[p];
      const int L = (int)ids.size();
      length[p] = (uint32_t)L;
      start[p] = pos;
      for (int q = 0; q < L; ++q) {
         cell[pos + q] = ids[q];
         dz[pos + q] = (float)dx[d];
         // target positions W-1 .. L-W, written when the cell is a normal cell (ratio
         // 1 / (Rratio * Rratio) = 1 without refinement)
         const bool target = q >= W - 1 && q <= L - W && cell_kind[ids[q]] == VACC_CELL_NORMAL;
         ratio[pos + q] = target ? 1.0f : 0.0f;
      }
      pos += (uint64_t)L;
   }
   return 0;
}
