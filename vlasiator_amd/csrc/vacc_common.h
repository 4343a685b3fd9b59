// vacc_common.h -- shared definitions of libvacc (host + device).
//
// Layout of a velocity block: 64 floats, in-block index i + 4*j + 16*k
// (common.h:512-516 of the reference).  GID = x + y*Nx + z*Nx*Ny
// (velocity_mesh_cpu.h:341-347).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vacc.h"

#define VACC_HD __host__ __device__ inline

namespace vacc {

constexpr int WID = 4;
constexpr int WID2 = 16;
constexpr int WID3 = 64;
constexpr uint32_t INVALID = 0xFFFFFFFFu;
constexpr int RING = 32;            // per-lane target ring (cells) of the remap kernel
constexpr int MAX_SPAN_CELLS = 22;  // (dv + 3|di| + 3|dj|)/dk bound that keeps the ring in range
// cells with a wider span go to the wide instance of the exact kernel: a 128-cell ring in
// workgroups of 64 lanes (32 KiB of LDS, as the 256-lane instance)
constexpr int RING_WIDE = 128;
constexpr int WG_WIDE = 64;
constexpr int MAX_SPAN_WIDE = RING_WIDE - 10;

// floor(g / d) for every 32-bit g without an integer division: (g + umulhi(g, m)) >> l, with
// l = ceil(log2 d) and m = ceil(2^(32+l) / d) - 2^32.  Exact: with M = m + 2^32 the product
// g M / 2^(32+l) exceeds g/d by less than g / 2^(32+l) < 2^-l <= 1/d, which cannot carry g/d
// past the next integer (its fraction is at most (d-1)/d).
struct FastDiv {
   uint32_t d, m, l;
};
inline FastDiv make_fastdiv(uint32_t d) {   // host side, d >= 1
   uint32_t l = 0;
   while (l < 32 && (1ull << l) < (uint64_t)d) ++l;
   const uint32_t k = 32 + l;   // ceil(2^k / d) = floor((2^k - 1) / d) + 1
   const uint64_t M = ((k == 64) ? ~0ull : ((1ull << k) - 1ull)) / d + 1ull;
   return FastDiv{d, (uint32_t)(M - (1ull << 32)), l};
}
VACC_HD uint32_t fdiv(uint32_t g, const FastDiv& f) {
#if defined(__HIP_DEVICE_COMPILE__)
   const uint32_t hi = __umulhi(g, f.m);
#else
   const uint32_t hi = (uint32_t)(((uint64_t)g * f.m) >> 32);
#endif
   return (uint32_t)(((uint64_t)g + hi) >> f.l);
}

// Block coordinates (x, y, z) of GID = x + y Nx + z Nx Ny.
struct GidSplit {
   FastDiv nx, nxy;
};
inline GidSplit make_gid_split(const uint32_t* len) { return GidSplit{make_fastdiv(len[0]), make_fastdiv(len[0] * len[1])}; }
VACC_HD void gid_xyz(uint32_t g, const GidSplit& s, uint32_t& x, uint32_t& y, uint32_t& z) {
   z = fdiv(g, s.nxy);
   const uint32_t r = g - z * s.nxy.d;
   y = fdiv(r, s.nx);
   x = r - y * s.nx.d;
}

// physical constants (common.h physicalconstants)
constexpr double MU_0 = 1.25663706e-6;
constexpr double PI = 3.14159265358979323846;

struct MeshInfo {
   uint32_t len[3];
   double vmin[3];
   double cell[3];      // cell size along each axis
};

VACC_HD double cell_size(const double vmin, const double vmax, const uint32_t len) {
   // velocity_mesh_parameters.cpp:153-163
   const double grid_size = vmax - vmin;
   const double block_size = grid_size / len;
   return block_size / WID;
}

// ---------------------------------------------------------------------------
// FP64 acceleration transform (cpu_acc_transform.cpp:66-153) as a 3x4 affine
// map, row-major {L00 L01 L02 t0 ; L10 ... ; L20 ...}.
// ---------------------------------------------------------------------------
struct Affine {
   double m[12];
};

VACC_HD void affine_point(const Affine& T, const double p[3], double o[3]) {
   for (int r = 0; r < 3; ++r)
      o[r] = ((T.m[4 * r] * p[0] + T.m[4 * r + 1] * p[1]) + T.m[4 * r + 2] * p[2]) + T.m[4 * r + 3];
}
VACC_HD void affine_vector(const Affine& T, const double p[3], double o[3]) {
   for (int r = 0; r < 3; ++r)
      o[r] = (T.m[4 * r] * p[0] + T.m[4 * r + 1] * p[1]) + T.m[4 * r + 2] * p[2];
}

// Rotation by angle th about the unit axis u (Rodrigues form).
VACC_HD void axis_rotation(const double u[3], double th, double R[9]) {
   const double s = sin(th), c = cos(th), oc = 1.0 - c;
   const double sx = s * u[0], sy = s * u[1], sz = s * u[2];
   const double ox = oc * u[0], oy = oc * u[1], oz = oc * u[2];
   R[0] = ox * u[0] + c;
   R[1] = ox * u[1] - sz;
   R[2] = ox * u[2] + sy;
   R[3] = ox * u[1] + sz;
   R[4] = oy * u[1] + c;
   R[5] = oy * u[2] - sx;
   R[6] = ox * u[2] - sy;
   R[7] = oy * u[2] + sx;
   R[8] = oz * u[2] + c;
}

struct TransformIn {
   double B[3];
   double dB[6];   // dBXdy dBXdz dBYdx dBYdz dBZdx dBZdy
   double V[3];
   double rhoq;
   double EgradPe[3];
   double dt;
};

VACC_HD Affine acceleration_transform(const TransformIn& in, double mass, double charge,
                                      double hall_min_rhoq, int ohm_gradpe) {
   const double n2 = in.B[0] * in.B[0] + in.B[1] * in.B[1] + in.B[2] * in.B[2];
   const double nrm = sqrt(n2);
   double u[3] = {in.B[0], in.B[1], in.B[2]};
   if (n2 > 0.0) {
      u[0] = in.B[0] / nrm;
      u[1] = in.B[1] / nrm;
      u[2] = in.B[2] / nrm;
   }
   const double B_mag = nrm + 1e-30;
   if (B_mag < 1e-28) {   // cpu_acc_transform.cpp:94-97
      u[0] = 0.0;
      u[1] = 0.0;
      u[2] = 1.0;
   }
   const double gyro_period = 2 * PI * mass / (charge * B_mag);
   const double eps = 1e10 * 2.2250738585072014e-308;   // 1e10 * numeric_limits<double>::min()
   const double rhoq = in.rhoq + eps;
   const double hall_rhoq = (rhoq <= hall_min_rhoq) ? hall_min_rhoq : rhoq;
   const double hall = 1.0 / (MU_0 * hall_rhoq);
   unsigned int substeps = (unsigned int)(fabs(in.dt) / fabs(gyro_period * (0.1 / 360.0)));
   if (substeps < 1) substeps = 1;
   const double rad = -(2.0 * PI * in.dt / gyro_period) / substeps;
   const double sdt = in.dt / substeps;
   double R[9];
   axis_rotation(u, rad, R);
   const double hx = hall * (in.dB[5] - in.dB[3]);
   const double hy = hall * (in.dB[1] - in.dB[4]);
   const double hz = hall * (in.dB[2] - in.dB[0]);
   Affine T = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}};
   for (unsigned int s = 0; s < substeps; ++s) {
      double p[3];
      affine_point(T, in.V, p);
      p[0] -= hx;
      p[1] -= hy;
      p[2] -= hz;
      for (int r = 0; r < 3; ++r) T.m[4 * r + 3] = -p[r] + T.m[4 * r + 3];
      Affine N;
      for (int r = 0; r < 3; ++r)
         for (int q = 0; q < 4; ++q)
            N.m[4 * r + q] = (R[3 * r] * T.m[q] + R[3 * r + 1] * T.m[4 + q]) + R[3 * r + 2] * T.m[8 + q];
      T = N;
      for (int r = 0; r < 3; ++r) T.m[4 * r + 3] = p[r] + T.m[4 * r + 3];
      if (ohm_gradpe > 0) {
         const double qm = fabs(charge) / mass;
         for (int r = 0; r < 3; ++r) T.m[4 * r + 3] = (qm * in.EgradPe[r]) * sdt + T.m[4 * r + 3];
      }
   }
   return T;
}

VACC_HD Affine affine_inverse(const Affine& T) {
   const double* a = T.m;
   const double c00 = a[5] * a[10] - a[6] * a[9];
   const double c10 = a[6] * a[8] - a[4] * a[10];
   const double c20 = a[4] * a[9] - a[5] * a[8];
   const double det = a[0] * c00 + a[1] * c10 + a[2] * c20;
   double L[9] = {c00 / det, (a[2] * a[9] - a[1] * a[10]) / det, (a[1] * a[6] - a[2] * a[5]) / det,
                  c10 / det, (a[0] * a[10] - a[2] * a[8]) / det, (a[2] * a[4] - a[0] * a[6]) / det,
                  c20 / det, (a[1] * a[8] - a[0] * a[9]) / det, (a[0] * a[5] - a[1] * a[4]) / det};
   Affine o;
   for (int r = 0; r < 3; ++r) {
      for (int q = 0; q < 3; ++q) o.m[4 * r + q] = L[3 * r + q];
      o.m[4 * r + 3] = -((L[3 * r] * a[3] + L[3 * r + 1] * a[7]) + L[3 * r + 2] * a[11]);
   }
   return o;
}

// line/plane intersection, cpu_acc_intersections.cpp:112-117
VACC_HD void line_plane(const double lp[3], const double ld[3], const double pp[3], const double pn[3],
                        double o[3]) {
   const double nom = (pn[0] * (pp[0] - lp[0]) + pn[1] * (pp[1] - lp[1])) + pn[2] * (pp[2] - lp[2]);
   const double dem = (pn[0] * ld[0] + pn[1] * ld[1]) + pn[2] * ld[2];
   const double t = nom / dem;
   for (int q = 0; q < 3; ++q) o[q] = lp[q] + t * ld[q];
}

// Lagrangian coordinates along `dim` of the four reference points -> (I, di, dj, dk).
VACC_HD void store_isect(const double p0[3], const double pi[3], const double pj[3], const double pk[3],
                         int dim, double* o4) {
   o4[0] = p0[dim];
   o4[1] = pi[dim] - p0[dim];
   o4[2] = pj[dim] - p0[dim];
   o4[3] = pk[dim] - p0[dim];
}

// SLICE-3D intersections (Zerroukat & Allen 2012, section 2.4) for the pass at position
// `pass` (0,1,2) of the map order, along dimension `dim` (cpu_acc_intersections.cpp:132-450).
VACC_HD void pass_intersections(const MeshInfo& m, const Affine& bwd, int pass, int dim, double* o4) {
   const double* dv = m.cell;
   const double* mn = m.vmin;
   double r[4][3];
   if (pass == 0) {
      // Lagrangian planes normal to bwd*e_dim, Euclidean lines along e_dim.
      double e[3] = {0, 0, 0}, pn[3], pp[3], base[3] = {0, 0, 0}, ld[3] = {0, 0, 0}, lp[3];
      e[dim] = 1.0;
      affine_vector(bwd, e, pn);
      base[dim] = mn[dim];
      affine_point(bwd, base, pp);
      ld[dim] = 1.0;
      for (int q = 0; q < 3; ++q) lp[q] = (q == dim) ? 0.0 : 0.5 * dv[q] + mn[q];
      line_plane(lp, ld, pp, pn, r[0]);
      for (int q = 0; q < 3; ++q) {
         double w[3] = {0, 0, 0}, a[3], b[3];
         w[q] = dv[q];
         if (q == dim) {
            double s[3];
            affine_vector(bwd, w, s);
            for (int t = 0; t < 3; ++t) { a[t] = lp[t]; b[t] = pp[t] + s[t]; }
         } else {
            for (int t = 0; t < 3; ++t) { a[t] = lp[t] + w[t]; b[t] = pp[t]; }
         }
         line_plane(a, ld, b, pn, r[q + 1]);
      }
   } else if (pass == 1) {
      // Euclidean planes normal to axis pa, Lagrangian lines along bwd*e_pa.
      const int pa = (dim == 0) ? 1 : (dim == 1) ? 2 : 0;
      double pn[3] = {0, 0, 0}, pp[3] = {0, 0, 0}, e[3] = {0, 0, 0}, ld[3], lpt[3], lp[3];
      pn[pa] = 1.0;
      pp[pa] = mn[pa] + dv[pa] * 0.5;
      e[pa] = 1.0;
      affine_vector(bwd, e, ld);
      for (int q = 0; q < 3; ++q) lpt[q] = (q == dim) ? mn[q] : 0.5 * dv[q] + mn[q];
      affine_point(bwd, lpt, lp);
      line_plane(lp, ld, pp, pn, r[0]);
      for (int q = 0; q < 3; ++q) {
         double w[3] = {0, 0, 0}, a[3], b[3];
         w[q] = dv[q];
         if (q == pa) {
            for (int t = 0; t < 3; ++t) { a[t] = lp[t]; b[t] = pp[t] + w[t]; }
         } else {
            double s[3];
            affine_vector(bwd, w, s);
            for (int t = 0; t < 3; ++t) { a[t] = lp[t] + s[t]; b[t] = pp[t]; }
         }
         line_plane(a, ld, b, pn, r[q + 1]);
      }
   } else {
      // Lagrangian images of cell-centre points.
      double base[3];
      for (int q = 0; q < 3; ++q) base[q] = (q == dim ? 0.0 : 0.5) * dv[q] + mn[q];
      affine_point(bwd, base, r[0]);
      for (int q = 0; q < 3; ++q) {
         double w[3];
         for (int t = 0; t < 3; ++t) {
            double frac = (t == dim ? 0.0 : 0.5);
            if (t == q) frac += 1.0;
            w[t] = frac * dv[t] + mn[t];
         }
         affine_point(bwd, w, r[q + 1]);
      }
   }
   store_isect(r[0], r[1], r[2], r[3], dim, o4);
}

VACC_HD int pass_dim(int map_order, int pass) {
   // 0: X-Y-Z, 1: Y-Z-X, 2: Z-X-Y (cpu_acc_semilag.cpp:98-129)
   return (map_order + pass) % 3;
}

// intersections[12] indexed by dimension (x,y,z) x (I,di,dj,dk)
VACC_HD void cell_intersections(const MeshInfo& m, const Affine& fwd, int map_order, double out[12]) {
   const Affine bwd = affine_inverse(fwd);
   for (int p = 0; p < 3; ++p) {
      const int d = pass_dim(map_order, p);
      pass_intersections(m, bwd, p, d, out + 4 * d);
   }
}

}  // namespace vacc
