/*
 * vacc_oracle.c -- CPU oracle for the velocity-space acceleration path.
 *
 * TEST INFRASTRUCTURE ONLY (see vacc_oracle.h).  Plain-C restatement of the
 * reference CPU algorithm of fmihpc/vlasiator built with VEC_FALLBACK_GENERIC,
 * VECL=8, Realf=float, Real=double.  Every vector operation of the reference is
 * restated per lane with the type promotion of vectorclass_fallback.h: a Vec
 * combined with a double scalar is evaluated in double and rounded to float
 * (vectorclass_fallback.h:98-226 and insert() :663), a Vec combined with an
 * int or float scalar is evaluated in float.  Build with -ffp-contract=off.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef VO_REALF_DOUBLE
/* FP64 build (libvacc_oracle64.so): Realf = double, as the reference built with DPF
 * (definitions.h: typedef double Realf).  Every FP32 lane value and rounding of the
 * restatement becomes FP64; the API takes double f arrays.  Used as the FP64 run that the
 * storage forms are measured against. */
#define float double
#define sqrtf sqrt
#define fabsf fabs
#endif

#include "vacc_oracle.h"

/* Vec<float> (op) double-scalar: computed in double, rounded into the float lane. */
static inline float dm(double c, float x) { return (float)(c * (double)x); }

double vo_cell_size(const vo_mesh* m, int dim) {
   /* velocity_mesh_parameters.cpp:153-163: gridSize/gridLength/blockLength */
   const double grid_size = m->vmax[dim] - m->vmin[dim];
   const double block_size = grid_size / m->len[dim];
   return block_size / VO_WID;
}

/* ------------------------------------------------------------------------ */
/* Face estimates, cpu_face_estimates.hpp                                    */
/* ------------------------------------------------------------------------ */

/* h4 left face value, cpu_face_estimates.hpp:160-167 */
static float h4_face(const float* v, int k) {
   float t = dm(-1.0, v[k - 2]);
   t = t + dm(7.0, v[k - 1]);
   t = t + dm(7.0, v[k]);
   t = t - dm(1.0, v[k + 1]);
   return (float)((1.0 / 12.0) * (double)t);
}

/* h8 left face value, cpu_face_estimates.hpp:45-56 */
static float h8_face(const float* v, int k) {
   float t = dm(-3.0, v[k - 4]);
   t = t + dm(29.0, v[k - 3]);
   t = t - dm(139.0, v[k - 2]);
   t = t + dm(533.0, v[k - 1]);
   t = t + dm(533.0, v[k]);
   t = t - dm(139.0, v[k + 1]);
   t = t + dm(29.0, v[k + 2]);
   t = t - dm(3.0, v[k + 3]);
   return (float)((1.0 / 840.0) * (double)t);
}

/* h7 left face derivative, cpu_face_estimates.hpp:68-78 */
static float h7_deriv(const float* v, int k) {
   float t = dm(9.0, v[k - 4]);
   t = t - dm(119.0, v[k - 3]);
   t = t + dm(889.0, v[k - 2]);
   t = t - dm(7175.0, v[k - 1]);
   t = t + dm(7175.0, v[k]);
   t = t - dm(889.0, v[k + 1]);
   t = t + dm(119.0, v[k + 2]);
   t = t - dm(9.0, v[k + 3]);
   return (float)((1.0 / 5040.0) * (double)t);
}

/* ------------------------------------------------------------------------ */
/* Slope limiters, cpu_slope_limiters.hpp                                    */
/* ------------------------------------------------------------------------ */

/* minmod / maxmod, cpu_slope_limiters.hpp:30-41 */
static inline float minmod(float s1, float s2) {
   const float slope = (fabsf(s1) < fabsf(s2)) ? s1 : s2;
   const float prod = s1 * s2;
   return (prod <= 0.0f) ? 0.0f : slope;
}
static inline float maxmod(float s1, float s2) {
   const float slope = (fabsf(s1) > fabsf(s2)) ? s1 : s2;
   const float prod = s1 * s2;
   return (prod <= 0.0f) ? 0.0f : slope;
}
/* superbee, cpu_slope_limiters.hpp:47-54 (slope_limiter() selects it, :97-101) */
static inline float superbee(float l, float m, float r) {
   const float a = r - m;
   const float b = m - l;
   const float s1 = minmod(a, 2.0f * b);
   const float s2 = minmod(2.0f * a, b);
   return maxmod(s1, s2);
}
/* abs/sign form, cpu_slope_limiters.hpp:113-118, applied on values scaled by 1/threshold
 * as in cpu_face_estimates.hpp:253-256 */
static inline void limiter_scaled(const float* v, int k, float thr, float* sabs, float* ssign) {
   const float scale = (float)(1.0 / (double)thr);
   const float l = v[k - 1] * scale;
   const float m = v[k] * scale;
   const float r = v[k + 1] * scale;
   const float s = superbee(l, m, r);
   *sabs = fabsf(s);
   *ssign = (s > 0.0f) ? 1.0f : -1.0f;
   *sabs = *sabs * thr;
}

/* ------------------------------------------------------------------------ */
/* Reconstructions                                                            */
/* ------------------------------------------------------------------------ */

/* PLM, cpu_1d_plm.hpp:38-49 */
void vo_plm_coeff(const float* v, int k, float thr, float a[2]) {
   const float scale = (float)(1.0 / (double)thr);
   const float d_cv = superbee(v[k - 1] * scale, v[k] * scale, v[k + 1] * scale) * thr;
   a[0] = v[k] - dm(0.5, d_cv);
   a[1] = dm(0.5, d_cv);
}

/* PPM face filter, cpu_face_estimates.hpp:278-323 (order h4) */
static void ppm_faces(const float* v, int k, float thr, float* pm, float* pp) {
   float fvl = h4_face(v, k);
   float fvr = h4_face(v, k + 1);
   float sabs, ssign;
   limiter_scaled(v, k, thr, &sabs, &ssign);
   if (sabs == 0.0f) {
      fvr = v[k];
      fvl = v[k];
   }
   {
      const float p = (v[k - 1] - fvl) * (fvl - v[k]);
      if (p < 0.0f) fvl = v[k] - dm(0.5, ssign) * sabs;
   }
   {
      const float p = (v[k + 1] - fvr) * (fvr - v[k]);
      if (p < 0.0f) fvr = v[k] + dm(0.5, ssign) * sabs;
   }
   *pm = fvl;
   *pp = fvr;
}

/* PPM coefficients, cpu_1d_ppm.hpp:35-54 */
void vo_ppm_coeff(const float* v, int k, float thr, float a[3]) {
   float m, p;
   ppm_faces(v, k, thr, &m, &p);
   const float one_sixth = (float)(1.0 / 6.0);
   {
      const float d = p - m;
      const float mid = v[k] - dm(0.5, m + p);
      const float lhs = d * mid;
      const float rhs = (d * d) * one_sixth;
      if (lhs > rhs) m = 3.0f * v[k] - 2.0f * p;
   }
   {
      const float d = p - m;
      const float mid = v[k] - dm(0.5, m + p);
      const float lhs = ((-d) * d) * one_sixth;
      const float rhs = d * mid;
      if (lhs > rhs) p = 3.0f * v[k] - 2.0f * m;
   }
   a[0] = m;
   a[1] = (dm(3.0, v[k]) - dm(2.0, m)) - p;
   a[2] = (m + p) - dm(2.0, v[k]);
}

/* PQM face values and derivatives, cpu_face_estimates.hpp:215-271 (order h8) */
static void pqm_faces(const float* v, int k, float thr, float* fvl, float* fvr, float* fdl, float* fdr) {
   float vl = h8_face(v, k), vr = h8_face(v, k + 1);
   float dl = h7_deriv(v, k), dr = h7_deriv(v, k + 1);
   float sabs, ssign;
   limiter_scaled(v, k, thr, &sabs, &ssign);
   if (sabs == 0.0f) {
      vr = v[k];
      vl = v[k];
      dl = 0.0f;
      dr = 0.0f;
   }
   {
      const float p = (v[k - 1] - vl) * (vl - v[k]);
      const float s = ssign * dl;
      if (p < 0.0f || s < 0.0f) {
         vl = v[k] - dm(0.5, ssign) * sabs;
         dl = ssign * sabs;
      }
   }
   {
      const float p = (v[k + 1] - vr) * (vr - v[k]);
      const float s = ssign * dr;
      if (p < 0.0f || s < 0.0f) {
         vr = v[k] + dm(0.5, ssign) * sabs;
         dr = ssign * sabs;
      }
   }
   *fvl = vl; *fvr = vr; *fdl = dl; *fdr = dr;
}

/* PQM monotonicity filter, cpu_1d_pqm.hpp:31-146 (White & Adcroft 2008 eqs 21, 23, 28, 29) */
static void pqm_monotonicity(const float* v, int k, float* pfvl, float* pfvr, float* pfdl, float* pfdr) {
   float fvl = *pfvl, fvr = *pfvr, fdl = *pfdl, fdr = *pfdr;
   const float vk = v[k];
   const float b0 = ((dm(60.0, vk) - dm(24.0, fvr)) - dm(36.0, fvl)) + dm(3.0, fdr - dm(3.0, fdl));
   const float b1 = (((dm(-360.0, vk) + dm(36.0, fdl)) - dm(24.0, fdr)) + dm(168.0, fvr)) + dm(192.0, fvl);
   const float b2 = (dm(360.0, vk) + dm(30.0, fdr - fdl)) - dm(180.0, fvl + fvr);
   const float val_to_sqrt = (b1 * b1) - (4.0f * b0) * b2;
   const float sqrt_val = (val_to_sqrt < 0.0f) ? (b1 + dm(200.0, b2)) : sqrtf(val_to_sqrt);
   const float two_b2 = 2.0f * b2;
   const float root1 = ((-b1) + sqrt_val) / two_b2;
   const float root2 = ((-b1) - sqrt_val) / two_b2;
   const float plm_slope_l = dm(2.0, vk - v[k - 1]);
   const float plm_slope_r = dm(2.0, v[k + 1] - vk);
   const float slope_sign = plm_slope_l + plm_slope_r;
   const float c0 = fdl;
   const float c1 = b0;
   const float c2 = (float)((double)b1 / 2.0);
   const float c3 = (float)((double)b2 / 3.0);
   const float r1p = c0 + root1 * (c1 + root1 * (c2 + root1 * c3));
   const float r2p = c0 + root2 * (c1 + root2 * (c2 + root2 * c3));
   const float root1_slope = (root1 >= 0.0f && root1 <= 1.0f) ? r1p : slope_sign;
   const float root2_slope = (root2 >= 0.0f && root2 <= 1.0f) ? r2p : slope_sign;
   const float t1 = root1_slope * slope_sign;
   const float t2 = root2_slope * slope_sign;
   if (t1 < 0.0f || t2 < 0.0f) {
      /* serialised per-lane fix (:93-138); scalar C++ with double literals */
      if (fabsf(plm_slope_l) <= fabsf(plm_slope_r)) {
         /* collapse to left edge (eq 21) */
         fdl = (float)((1.0 / 3.0) * (((double)(10.0f * vk) - 2.0 * (double)fvr) - 8.0 * (double)fvl));
         fdr = (float)((-10.0 * (double)vk + 6.0 * (double)fvr) + 4.0 * (double)fvl);
         if (slope_sign * fdl < 0.0f) {
            fdl = 0.0f;
            fvr = 5.0f * vk - 4.0f * fvl;
            fdr = 20.0f * (vk - fvl);
         } else if (slope_sign * fdr < 0.0f) {
            fdr = 0.0f;
            fvl = (float)(0.5 * (double)(5.0f * vk - 3.0f * fvr));
            fdl = (float)((10.0 / 3.0) * (double)(-vk + fvr));
         }
      } else {
         /* collapse to right edge (eq 21) */
         fdl = (float)((10.0 * (double)vk - 6.0 * (double)fvl) - 4.0 * (double)fvr);
         fdr = (float)((1.0 / 3.0) * ((-10.0 * (double)vk + (double)(2.0f * fvl)) + (double)(8.0f * fvr)));
         if (slope_sign * fdl < 0.0f) {
            fdl = 0.0f;
            fvr = (float)(0.5 * (double)(5.0f * vk - 3.0f * fvl));
            fdr = (float)((10.0 / 3.0) * (double)(vk - fvl));
         } else if (slope_sign * fdr < 0.0f) {
            fdr = 0.0f;
            fvl = 5.0f * vk - 4.0f * fvr;
            fdl = (float)(20.0 * (double)(-vk + fvr));
         }
      }
   }
   *pfvl = fvl; *pfvr = fvr; *pfdl = fdl; *pfdr = fdr;
}

/* PQM coefficients, cpu_1d_pqm.hpp:157-173 */
void vo_pqm_coeff(const float* v, int k, float thr, float a[5]) {
   float fvl, fvr, fdl, fdr;
   pqm_faces(v, k, thr, &fvl, &fvr, &fdl, &fdr);
   pqm_monotonicity(v, k, &fvl, &fvr, &fdl, &fdr);
   const float vk = v[k];
   a[0] = fvl;
   a[1] = (float)((double)fdl / 2.0);
   a[2] = ((dm(10.0, vk) - dm(4.0, fvr)) - dm(6.0, fvl)) + dm(0.5, fdr - 3.0f * fdl);
   a[3] = (((dm(-15.0, vk) + dm(1.5, fdl)) - fdr) + dm(7.0, fvr)) + 8.0f * fvl;
   a[4] = (dm(6.0, vk) + dm(0.5, fdr - fdl)) - dm(3.0, fvl + fvr);
}

void vo_coeff_batch(int method, const float* stencils, int n, float thr, float* out) {
   for (int i = 0; i < n; ++i) {
      const float* s = stencils + 9 * i;
      if (method == VO_PLM) vo_plm_coeff(s, 4, thr, out + 2 * i);
      else if (method == VO_PPM) vo_ppm_coeff(s, 4, thr, out + 3 * i);
      else vo_pqm_coeff(s, 4, thr, out + 5 * i);
   }
}

/* ------------------------------------------------------------------------ */
/* Column sort, cpu_acc_sort_blocks.cpp:46-155                                */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t key; uint32_t gid; uint32_t lid; } key_t3;

static int cmp_key(const void* a, const void* b) {
   const uint32_t ka = ((const key_t3*)a)->key, kb = ((const key_t3*)b)->key;
   return (ka < kb) ? -1 : (ka > kb) ? 1 : 0;
}

static uint32_t mapped_key(const vo_mesh* m, uint32_t gid, int dim) {
   const uint32_t nx = m->len[0], ny = m->len[1], nz = m->len[2];
   const uint32_t x = gid % nx, y = (gid / nx) % ny, z = gid / (nx * ny);
   switch (dim) {
      case 0: return gid;                                   /* :68-71 */
      case 1: return gid - (x + y * nx) + y + x * ny;       /* :84-89 */
      default: return z + y * nz + x * ny * nz;             /* :106-109 */
   }
}

/* Sorts and splits into columns/sets; also returns the input LID of each sorted block. */
static void sort_columns_impl(const vo_mesh* m, const uint32_t* gids, uint32_t n, int dim,
                              uint32_t* sorted, uint32_t* sorted_lid,
                              uint32_t* col_off, uint32_t* col_len, uint32_t* n_cols,
                              uint32_t* set_col_off, uint32_t* set_ncols, uint32_t* n_sets) {
   key_t3* kp = (key_t3*)malloc(sizeof(key_t3) * (n ? n : 1));
   for (uint32_t i = 0; i < n; ++i) {
      kp[i].key = mapped_key(m, gids[i], dim);
      kp[i].gid = gids[i];
      kp[i].lid = i;
   }
   qsort(kp, n, sizeof(key_t3), cmp_key);   /* keys are unique, so any sort is std::sort's order */
   uint32_t nc = 0, ns = 0;
   col_off[nc] = 0;
   set_col_off[ns] = 0;
   uint32_t prev_col = 0, prev_dim = 0;
   const uint32_t L = m->len[dim];
   for (uint32_t i = 0; i < n; ++i) {
      const uint32_t column_id = kp[i].key / L;
      const uint32_t dimension_id = kp[i].key % L;
      sorted[i] = kp[i].gid;
      if (sorted_lid) sorted_lid[i] = kp[i].lid;
      if (i > 0 && (column_id != prev_col || dimension_id != prev_dim + 1)) {
         col_off[nc + 1] = i;
         col_len[nc] = col_off[nc + 1] - col_off[nc];
         nc++;
         if (column_id != prev_col) {
            set_col_off[ns + 1] = nc;
            set_ncols[ns] = set_col_off[ns + 1] - set_col_off[ns];
            ns++;
         }
      }
      prev_col = column_id;
      prev_dim = dimension_id;
   }
   if (n > 0) {
      col_len[nc] = n - col_off[nc];
      nc++;
      set_ncols[ns] = nc - set_col_off[ns];
      ns++;
   }
   *n_cols = nc;
   *n_sets = ns;
   free(kp);
}

int vo_sort_columns(const vo_mesh* m, const uint32_t* gids, uint32_t n, int dim,
                    uint32_t* sorted, uint32_t* col_off, uint32_t* col_len, uint32_t* n_cols,
                    uint32_t* set_col_off, uint32_t* set_ncols, uint32_t* n_sets) {
   sort_columns_impl(m, gids, n, dim, sorted, NULL, col_off, col_len, n_cols, set_col_off, set_ncols, n_sets);
   return 0;
}

/* ------------------------------------------------------------------------ */
/* map_1d, cpu_acc_map.cpp:121-636                                            */
/* ------------------------------------------------------------------------ */

static inline void block_indices(const vo_mesh* m, uint32_t gid, uint32_t idx[3]) {
   /* velocity_mesh_cpu.h:361-369 */
   idx[0] = gid % m->len[0];
   idx[1] = (gid / m->len[0]) % m->len[1];
   idx[2] = gid / (m->len[0] * m->len[1]);
}
static inline void swap_indices(uint32_t idx[3], int dim) {
   /* cpu_acc_map.cpp:85-106 */
   uint32_t t;
   if (dim == 0) { t = idx[2]; idx[2] = idx[0]; idx[0] = t; }
   else if (dim == 1) { t = idx[2]; idx[2] = idx[1]; idx[1] = t; }
}

static void reconstruct(int method, const float* v, int k, float thr, float* a) {
   if (method == VO_PLM) vo_plm_coeff(v, k, thr, a);
   else if (method == VO_PPM) vo_ppm_coeff(v, k, thr, a);
   else vo_pqm_coeff(v, k, thr, a);
}

static inline float integrand(int method, float x, const float* a) {
   /* cpu_acc_map.cpp:590-602 */
   if (method == VO_PLM) return x * (a[0] + x * a[1]);
   if (method == VO_PPM) return x * (a[0] + x * (a[1] + x * a[2]));
   return x * (a[0] + x * (a[1] + x * (a[2] + x * (a[3] + x * a[4]))));
}

int vo_map_1d(const vo_mesh* m, uint32_t n_in, const uint32_t* gids_in, const float* f_in,
              double I_d, double di_d, double dj_d, double dk_d, int dim, int method,
              double min_value, int wall_margin,
              uint32_t cap_out, uint32_t* n_out, uint32_t* gids_out, float* f_out, int* wall_flag) {
   /* :128-131 conversion to Realf, then :151-199 swaps */
   float I = (float)I_d, di = (float)di_d, dj = (float)dj_d, dk = (float)dk_d;
   *n_out = 0;
   if (n_in == 0) return 0;   /* :143-145 */
   const float dv = (float)vo_cell_size(m, dim);
   const float v_min = (float)m->vmin[dim];
   const int max_v_length = (int)m->len[dim];
   const uint32_t nx = m->len[0], ny = m->len[1];
   uint32_t bid[3], cid[3];
   if (dim == 0) {
      float t = di; di = dk; dk = t;
      bid[0] = nx * ny; bid[1] = nx; bid[2] = 1;
      cid[0] = VO_WID2; cid[1] = VO_WID; cid[2] = 1;
   } else if (dim == 1) {
      float t = dj; dj = dk; dk = t;
      bid[0] = 1; bid[1] = nx * ny; bid[2] = nx;
      cid[0] = 1; cid[1] = VO_WID2; cid[2] = VO_WID;
   } else {
      bid[0] = 1; bid[1] = nx; bid[2] = nx * ny;
      cid[0] = 1; cid[1] = VO_WID; cid[2] = VO_WID2;
   }
   const double i_dv = 1.0 / (double)dv;   /* :201, Real */
   const float thr = (float)min_value;     /* Realf threshold argument */

   uint32_t* sorted = (uint32_t*)malloc(sizeof(uint32_t) * n_in);
   uint32_t* sorted_lid = (uint32_t*)malloc(sizeof(uint32_t) * n_in);
   uint32_t* col_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_in + 1));
   uint32_t* col_len = (uint32_t*)malloc(sizeof(uint32_t) * (n_in + 1));
   uint32_t* set_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_in + 1));
   uint32_t* set_nc = (uint32_t*)malloc(sizeof(uint32_t) * (n_in + 1));
   uint32_t ncols, nsets;
   sort_columns_impl(m, gids_in, n_in, dim, sorted, sorted_lid, col_off, col_len, &ncols, set_off, set_nc, &nsets);

   int* colMin = (int*)malloc(sizeof(int) * ncols);
   int* colMax = (int*)malloc(sizeof(int) * ncols);
   /* per-lane padded column values: 16 pencils x 4*(n+2) */
   uint32_t maxlen = 0;
   for (uint32_t c = 0; c < ncols; ++c) if (col_len[c] > maxlen) maxlen = col_len[c];
   const uint32_t stride = VO_WID * (maxlen + 2);
   float* vals = (float*)malloc(sizeof(float) * 16 * stride);

   float* target[VO_MAX_BLOCKS_PER_DIM];
   uint32_t nout = 0;
   int overflow = 0;

   for (uint32_t s = 0; s < nsets; ++s) {
      unsigned char isTarget[VO_MAX_BLOCKS_PER_DIM], isSource[VO_MAX_BLOCKS_PER_DIM];
      memset(isTarget, 0, sizeof(isTarget));
      memset(isSource, 0, sizeof(isSource));
      for (int b = 0; b < VO_MAX_BLOCKS_PER_DIM; ++b) target[b] = NULL;

      uint32_t sfi[3];
      block_indices(m, sorted[col_off[set_off[s]]], sfi);
      swap_indices(sfi, dim);
      /* :258-288 corner intersections (float) */
      const float x0 = (float)(sfi[0] * VO_WID + 0), x3 = (float)(sfi[0] * VO_WID + VO_WID - 1);
      const float y0 = (float)(sfi[1] * VO_WID + 0), y3 = (float)(sfi[1] * VO_WID + VO_WID - 1);
      const float c00 = (I + x0 * di) + y0 * dj;
      const float c03 = (I + x0 * di) + y3 * dj;
      const float c30 = (I + x3 * di) + y0 * dj;
      const float c33 = (I + x3 * di) + y3 * dj;
      float max_iMin = c00;
      max_iMin = (max_iMin < c03) ? c03 : max_iMin;   /* std::max(a,b) = (a<b)?b:a */
      max_iMin = (max_iMin < c30) ? c30 : max_iMin;
      max_iMin = (max_iMin < c33) ? c33 : max_iMin;
      float min_iMin = c00;
      min_iMin = (c03 < min_iMin) ? c03 : min_iMin;   /* std::min(a,b) = (b<a)?b:a */
      min_iMin = (c30 < min_iMin) ? c30 : min_iMin;
      min_iMin = (c33 < min_iMin) ? c33 : min_iMin;

      for (uint32_t c = set_off[s]; c < set_off[s] + set_nc[s]; ++c) {
         uint32_t fi[3], li[3];
         block_indices(m, sorted[col_off[c]], fi);
         block_indices(m, sorted[col_off[c] + col_len[c] - 1], li);
         swap_indices(fi, dim);
         swap_indices(li, dim);
         /* :307-318 */
         const float firstBlockMinV = (float)(VO_WID * fi[2]) * dv + v_min;
         const float lastBlockMaxV = (float)(VO_WID * (li[2] + 1)) * dv + v_min;
         const int firstBlock_gk = (int)((firstBlockMinV - max_iMin) / dk);
         const int lastBlock_gk = (int)((lastBlockMaxV - min_iMin) / dk);
         int fk = firstBlock_gk / VO_WID;
         int lk = lastBlock_gk / VO_WID;
         /* :320-324 */
         fk = (fk >= 0) ? fk : 0;
         fk = (fk < max_v_length) ? fk : max_v_length - 1;
         lk = (lk >= 0) ? lk : 0;
         lk = (lk < max_v_length) ? lk : max_v_length - 1;
         if (fk < wall_margin || fk >= max_v_length - wall_margin ||
             lk < wall_margin || lk >= max_v_length - wall_margin) {
            if (wall_flag) *wall_flag = 1;   /* :325-338 bailout, continue with clamped extents */
         }
         for (uint32_t b = fi[2]; b <= li[2]; ++b) isSource[b] = 1;
         for (int b = fk; b <= lk; ++b) isTarget[b] = 1;
         colMin[c] = fk;
         colMax[c] = lk;
      }
      /* :357-389 targets = union of column target ranges; created zeroed */
      for (int b = 0; b < VO_MAX_BLOCKS_PER_DIM; ++b) {
         if (!isTarget[b]) continue;
         if (nout < cap_out) {
            gids_out[nout] = sfi[0] * bid[0] + sfi[1] * bid[1] + (uint32_t)b * bid[2];
            target[b] = f_out + (size_t)nout * VO_WID3;
            memset(target[b], 0, sizeof(float) * VO_WID3);
         } else {
            overflow = 1;
         }
         nout++;
      }
      if (overflow) continue;

      /* :393-631 per column mapping */
      for (uint32_t c = set_off[s]; c < set_off[s] + set_nc[s]; ++c) {
         const uint32_t nb = col_len[c];
         const uint32_t cs = VO_WID * (nb + 2);
         /* cpu_acc_load_blocks.cpp:28-566: transposed load with one zero block each side */
         for (int p = 0; p < 16; ++p) {
            float* pv = vals + p * stride;
            for (uint32_t t = 0; t < cs; ++t) pv[t] = 0.0f;
            const int ip = p & 3, jp = p >> 2;
            for (uint32_t b = 0; b < nb; ++b) {
               const float* src = f_in + (size_t)sorted_lid[col_off[c] + b] * VO_WID3;
               for (int kk = 0; kk < VO_WID; ++kk)
                  pv[VO_WID * (b + 1) + kk] = src[ip * cid[0] + jp * cid[1] + kk * cid[2]];
            }
         }
         uint32_t bb[3];
         block_indices(m, sorted[col_off[c]], bb);
         swap_indices(bb, dim);
         for (int j = 0; j < VO_WID; j += 2) {      /* j-vector loop, VECL=8 */
            float Imin[8], lvr[8];
            int gkr[8];
            const float v_r0 = (float)(VO_WID * bb[2]) * dv + v_min;
            for (int L = 0; L < 8; ++L) {
               const int ip = L & 3, jp = j + (L >> 2);
               const float X = (float)(bb[0] * VO_WID) + (float)ip;
               const float Y = (float)(bb[1] * VO_WID) + (float)jp;
               Imin[L] = (I + X * di) + Y * dj;           /* :493-496 */
               lvr[L] = (v_r0 - Imin[L]) / dk;            /* :503 */
               gkr[L] = (int)lvr[L];                      /* :507 */
            }
            int minIdx = 0, maxIdx = 0;                    /* :515-529 */
            {
               double maxV = DBL_MIN, minV = DBL_MAX;
               for (int L = 0; L < 8; ++L) {
                  if ((double)lvr[L] > maxV) { maxV = lvr[L]; maxIdx = L; }
                  if ((double)lvr[L] < minV) { minV = lvr[L]; minIdx = L; }
               }
            }
            float v_r = v_r0;
            for (uint32_t k = 0; k < VO_WID * nb; ++k) {
               float a[8][5];
               for (int L = 0; L < 8; ++L) {
                  const int p = (L & 3) + 4 * (j + (L >> 2));
                  reconstruct(method, vals + p * stride, (int)k + VO_WID, thr, a[L]);
               }
               float tdr[8];
               for (int L = 0; L < 8; ++L) tdr[L] = 0.0f;
               const float v_l = v_r;
               v_r = v_r + dv;
               int gkl[8];
               for (int L = 0; L < 8; ++L) {
                  gkl[L] = gkr[L];
                  gkr[L] = (int)((v_r - Imin[L]) / dk);   /* :561-563 */
               }
               const int lo = colMin[c] * VO_WID;
               const int hi = (colMax[c] + 1) * VO_WID - 1;
               const int minGk = (gkl[minIdx] > lo) ? gkl[minIdx] : lo;   /* :568 */
               const int maxGk = (gkr[maxIdx] < hi) ? gkr[maxIdx] : hi;   /* :569 */
               for (int gk = minGk; gk <= maxGk; ++gk) {
                  const int blockK = gk / VO_WID;
                  const int gkm = gk - blockK * VO_WID;
                  float* tb = target[blockK];
                  const float gkdk = (float)(gk + 1) * dk;
                  for (int L = 0; L < 8; ++L) {
                     const int ip = L & 3, jp = j + (L >> 2);
                     const float y = gkdk + Imin[L];
                     const float mx = (y > v_l) ? y : v_l;
                     const float mn = (mx < v_r) ? mx : v_r;
                     const float vn = (float)((double)(mn - v_l) * i_dv);   /* :585 */
                     const float tdl = tdr[L];
                     tdr[L] = integrand(method, vn, a[L]);
                     const float tval = tdr[L] - tdl;
                     const uint32_t tcell = ip * cid[0] + jp * cid[1] + gkm * cid[2];
                     tb[tcell] = tb[tcell] + tval;                          /* :609-625 */
                  }
               }
            }
         }
      }
   }
   *n_out = nout;
   free(sorted); free(sorted_lid); free(col_off); free(col_len); free(set_off); free(set_nc);
   free(colMin); free(colMax); free(vals);
   return overflow ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Cell container + 3-D step (cpu_acc_semilag.cpp:92-130)                     */
/* ------------------------------------------------------------------------ */

int vo_cell_init(vo_cell* c, uint32_t n, const uint32_t* gids, const float* f) {
   c->n = n;
   c->cap = n > 16 ? n : 16;
   c->gids = (uint32_t*)malloc(sizeof(uint32_t) * c->cap);
   c->f = (float*)malloc(sizeof(float) * VO_WID3 * (size_t)c->cap);
   if (!c->gids || !c->f) return -1;
   if (n) {
      memcpy(c->gids, gids, sizeof(uint32_t) * n);
      memcpy(c->f, f, sizeof(float) * VO_WID3 * (size_t)n);
   }
   return 0;
}

void vo_cell_free(vo_cell* c) {
   free(c->gids);
   free(c->f);
   c->gids = NULL;
   c->f = NULL;
   c->n = c->cap = 0;
}

static int cell_map(const vo_mesh* m, vo_cell* c, const double* isect4, int dim, int method,
                    double min_value, int wall_margin, int* wall_flag) {
   uint32_t cap = c->n + c->n / 2 + 64;
   for (;;) {
      uint32_t* g = (uint32_t*)malloc(sizeof(uint32_t) * cap);
      float* f = (float*)malloc(sizeof(float) * VO_WID3 * (size_t)cap);
      uint32_t nout = 0;
      int rc = vo_map_1d(m, c->n, c->gids, c->f, isect4[0], isect4[1], isect4[2], isect4[3], dim, method,
                         min_value, wall_margin, cap, &nout, g, f, wall_flag);
      if (rc == 0) {
         free(c->gids);
         free(c->f);
         c->gids = g;
         c->f = f;
         c->n = nout;
         c->cap = cap;
         return 0;
      }
      free(g);
      free(f);
      cap = nout + 64;
   }
}

int vo_accelerate_cell(const vo_mesh* m, vo_cell* c, const double intersections[12], int map_order,
                       int method, double min_value, int wall_margin, int* wall_flag) {
   static const int orders[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
   if (map_order < 0 || map_order > 2) return -2;
   for (int p = 0; p < 3; ++p) {
      const int dim = orders[map_order][p];
      cell_map(m, c, intersections + 4 * dim, dim, method, min_value, wall_margin, wall_flag);
   }
   return 0;
}

/* ------------------------------------------------------------------------ */
/* Sparsity / block adjust, spatial_cell_cpu.cpp:179-316                     */
/* ------------------------------------------------------------------------ */

void vo_content_flags(uint32_t n, const float* f, double min_value, uint8_t* flags) {
   for (uint32_t b = 0; b < n; ++b) {
      uint8_t has = 0;
      for (int i = 0; i < VO_WID3; ++i) {
         if ((double)f[(size_t)b * VO_WID3 + i] >= min_value) { has = 1; break; }
      }
      flags[b] = has;
   }
}

static int cmp_u32(const void* a, const void* b) {
   const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
   return (x < y) ? -1 : (x > y) ? 1 : 0;
}

int vo_adjust_spatial(const vo_mesh* m, uint32_t n_in, const uint32_t* gids, const float* f,
                      double min_value, int add_width, uint32_t n_extra, const uint32_t* extra_gids,
                      uint32_t cap, uint32_t* n_out, uint32_t* gids_out, float* f_out, double* rho_loss) {
   const uint32_t nx = m->len[0], ny = m->len[1], nz = m->len[2];
   const size_t ntot = (size_t)nx * ny * nz;
   /* dense maps over the mesh: keep flag and LID+1 of existing blocks */
   uint8_t* keep = (uint8_t*)calloc(ntot, 1);
   uint32_t* lid1 = (uint32_t*)calloc(ntot, sizeof(uint32_t));
   uint8_t* content = (uint8_t*)malloc(n_in ? n_in : 1);
   vo_content_flags(n_in, f, min_value, content);
   for (uint32_t b = 0; b < n_in; ++b) lid1[gids[b]] = b + 1;
   /* :190-210 content blocks and their (+-addWidthV)^3 velocity neighbours */
   for (uint32_t b = 0; b < n_in; ++b) {
      if (!content[b]) continue;
      uint32_t idx[3];
      block_indices(m, gids[b], idx);
      keep[gids[b]] = 1;
      for (int ox = -add_width; ox <= add_width; ++ox)
         for (int oy = -add_width; oy <= add_width; ++oy)
            for (int oz = -add_width; oz <= add_width; ++oz) {
               /* unsigned wrap of negative indices -> out of range -> invalid (velocity_mesh_cpu.h:327-347) */
               const uint32_t x = idx[0] + (uint32_t)ox, y = idx[1] + (uint32_t)oy, z = idx[2] + (uint32_t)oz;
               if (x >= nx || y >= ny || z >= nz) continue;
               keep[x + y * nx + (size_t)z * nx * ny] = 1;
            }
   }
   /* :212-222 content blocks of the spatial neighbours, kept as they are (no dilation) */
   for (uint32_t i = 0; i < n_extra; ++i)
      if ((size_t)extra_gids[i] < ntot) keep[extra_gids[i]] = 1;
   /* :227-267 remove no-content blocks not kept, iterating the no-content list in reverse */
   const double DV3 = vo_cell_size(m, 0) * vo_cell_size(m, 1) * vo_cell_size(m, 2);
   double loss = 0.0;
   for (int64_t b = (int64_t)n_in - 1; b >= 0; --b) {
      if (content[b]) continue;
      if (keep[gids[b]]) continue;
      double sum = 0.0;
      for (int i = 0; i < VO_WID3; ++i) sum += f[(size_t)b * VO_WID3 + i];
      loss += DV3 * sum;
   }
   /* :270-272 add every kept block; output sorted by GID */
   uint32_t nk = 0;
   for (size_t g = 0; g < ntot; ++g) if (keep[g]) nk++;
   *n_out = nk;
   if (rho_loss) *rho_loss = loss;
   if (nk > cap) {
      free(keep); free(lid1); free(content);
      return -1;
   }
   uint32_t o = 0;
   for (size_t g = 0; g < ntot; ++g) {
      if (!keep[g]) continue;
      gids_out[o] = (uint32_t)g;
      if (lid1[g]) memcpy(f_out + (size_t)o * VO_WID3, f + (size_t)(lid1[g] - 1) * VO_WID3, sizeof(float) * VO_WID3);
      else memset(f_out + (size_t)o * VO_WID3, 0, sizeof(float) * VO_WID3);
      o++;
   }
   free(keep); free(lid1); free(content);
   (void)cmp_u32;
   return 0;
}

/* spatial_cell_cpu.cpp:699-727 */
double vo_sparse_min_value(int algorithm, double min_value, double bulk1, double bulk2, double min1, double min2,
                           double rho, uint32_t nblocks) {
   if (algorithm != 1 && algorithm != 2) return min_value;
   const double k = (min2 - min1) / (bulk2 - bulk1);
   const double b = min1 - k * bulk1;
   const double x = (algorithm == 1) ? rho : (double)nblocks;
   const double y = k * x + b;
   if (y < min1) return min1;
   if (y > min2) return min2;
   return y;
}

/* block_adjust_cpu.cpp:96-113 */
void vo_conserve_mass(uint32_t n_pre, const float* f_pre, uint32_t n_post, float* f_post) {
   double pre = 0.0, post = 0.0;
   for (size_t i = 0; i < (size_t)n_pre * VO_WID3; ++i) pre += f_pre[i];
   for (size_t i = 0; i < (size_t)n_post * VO_WID3; ++i) post += f_post[i];
   if (post != 0.0)
      for (size_t i = 0; i < (size_t)n_post * VO_WID3; ++i) f_post[i] = (float)((double)f_post[i] * (pre / post));
}

int vo_adjust(const vo_mesh* m, uint32_t n_in, const uint32_t* gids, const float* f,
              double min_value, int add_width,
              uint32_t cap, uint32_t* n_out, uint32_t* gids_out, float* f_out, double* rho_loss) {
   return vo_adjust_spatial(m, n_in, gids, f, min_value, add_width, 0, NULL, cap, n_out, gids_out, f_out, rho_loss);
}

/* ------------------------------------------------------------------------ */
/* Moments, arch_moments.h:83-108                                            */
/* ------------------------------------------------------------------------ */

void vo_moments(const vo_mesh* m, uint32_t n, const uint32_t* gids, const float* f, double out[4]) {
   const double dvx = vo_cell_size(m, 0), dvy = vo_cell_size(m, 1), dvz = vo_cell_size(m, 2);
   const double DV3 = dvx * dvy * dvz;
   double s[4] = {0, 0, 0, 0};
   for (uint32_t b = 0; b < n; ++b) {
      uint32_t idx[3];
      block_indices(m, gids[b], idx);
      const double bx = m->vmin[0] + idx[0] * (dvx * VO_WID);
      const double by = m->vmin[1] + idx[1] * (dvy * VO_WID);
      const double bz = m->vmin[2] + idx[2] * (dvz * VO_WID);
      for (int k = 0; k < VO_WID; ++k)
         for (int j = 0; j < VO_WID; ++j)
            for (int i = 0; i < VO_WID; ++i) {
               const double VX = bx + (i + 0.5) * dvx;
               const double VY = by + (j + 0.5) * dvy;
               const double VZ = bz + (k + 0.5) * dvz;
               const double a = f[(size_t)b * VO_WID3 + k * VO_WID2 + j * VO_WID + i];
               s[0] += a * DV3;
               s[1] += a * VX * DV3;
               s[2] += a * VY * DV3;
               s[3] += a * VZ * DV3;
            }
   }
   out[0] = s[0]; out[1] = s[1]; out[2] = s[2]; out[3] = s[3];
}

/* arch_moments.h:123-150 */
void vo_pressure(const vo_mesh* m, uint32_t n, const uint32_t* gids, const float* f, const double V[3],
                 double mass, double out[6]) {
   const double dv[3] = {vo_cell_size(m, 0), vo_cell_size(m, 1), vo_cell_size(m, 2)};
   const double DV3 = dv[0] * dv[1] * dv[2];
   double s[6] = {0, 0, 0, 0, 0, 0};
   for (uint32_t b = 0; b < n; ++b) {
      uint32_t idx[3];
      block_indices(m, gids[b], idx);
      for (int k = 0; k < VO_WID; ++k)
         for (int j = 0; j < VO_WID; ++j)
            for (int i = 0; i < VO_WID; ++i) {
               const double a = f[(size_t)b * VO_WID3 + i + VO_WID * j + VO_WID2 * k];
               const double dx = (m->vmin[0] + (idx[0] * VO_WID + i + 0.5) * dv[0]) - V[0];
               const double dy = (m->vmin[1] + (idx[1] * VO_WID + j + 0.5) * dv[1]) - V[1];
               const double dz = (m->vmin[2] + (idx[2] * VO_WID + k + 0.5) * dv[2]) - V[2];
               s[0] += a * dx * dx * DV3;
               s[1] += a * dy * dy * DV3;
               s[2] += a * dz * dz * DV3;
               s[3] += a * dy * dz * DV3;
               s[4] += a * dx * dz * DV3;
               s[5] += a * dx * dy * DV3;
            }
   }
   for (int q = 0; q < 6; ++q) out[q] = mass * s[q];
}

/* ------------------------------------------------------------------------ */
/* Transform (cpu_acc_transform.cpp:66-153) and intersections                */
/* (cpu_acc_intersections.cpp:44-450), FP64                                   */
/* ------------------------------------------------------------------------ */

#define VO_MU_0 1.25663706e-6        /* common.h physicalconstants */
#define VO_PI 3.14159265358979323846

/* Rodrigues rotation about unit axis u by angle th (right-handed). */
static void rotation_matrix(const double u[3], double th, double R[9]) {
   const double s = sin(th), c = cos(th);
   const double su0 = s * u[0], su1 = s * u[1], su2 = s * u[2];
   const double cu0 = (1.0 - c) * u[0], cu1 = (1.0 - c) * u[1], cu2 = (1.0 - c) * u[2];
   R[0 * 3 + 1] = cu0 * u[1] - su2;
   R[1 * 3 + 0] = cu0 * u[1] + su2;
   R[0 * 3 + 2] = cu0 * u[2] + su1;
   R[2 * 3 + 0] = cu0 * u[2] - su1;
   R[1 * 3 + 2] = cu1 * u[2] - su0;
   R[2 * 3 + 1] = cu1 * u[2] + su0;
   R[0] = cu0 * u[0] + c;
   R[4] = cu1 * u[1] + c;
   R[8] = cu2 * u[2] + c;
}

static void affine_apply(const double T[12], const double p[3], double out[3]) {
   for (int i = 0; i < 3; ++i)
      out[i] = ((T[i * 4 + 0] * p[0] + T[i * 4 + 1] * p[1]) + T[i * 4 + 2] * p[2]) + T[i * 4 + 3];
}
static void linear_apply(const double T[12], const double p[3], double out[3]) {
   for (int i = 0; i < 3; ++i)
      out[i] = (T[i * 4 + 0] * p[0] + T[i * 4 + 1] * p[1]) + T[i * 4 + 2] * p[2];
}

void vo_transform(const vo_fields* fl, double mass, double charge, double dt,
                  double hall_min_rhoq, int ohm_gradpe, double out[12]) {
   const double Bx = fl->B[0], By = fl->B[1], Bz = fl->B[2];
   const double n2 = Bx * Bx + By * By + Bz * Bz;
   const double nrm = sqrt(n2);
   double u[3] = {Bx, By, Bz};
   if (n2 > 0.0) { u[0] = Bx / nrm; u[1] = By / nrm; u[2] = Bz / nrm; }
   const double B_mag = nrm + 1e-30;
   if (B_mag < 1e-28) { u[0] = 0; u[1] = 0; u[2] = 1; }   /* :94-97 */
   const double gyro_period = 2 * VO_PI * mass / (charge * B_mag);
   const double EPSILON = 1e10 * DBL_MIN;
   const double rhoq = fl->rhoq + EPSILON;
   const double hallRhoq = (rhoq <= hall_min_rhoq) ? hall_min_rhoq : rhoq;
   const double hallPrefactor = 1.0 / (VO_MU_0 * hallRhoq);
   const double dBXdy = fl->dB[0], dBXdz = fl->dB[1], dBYdx = fl->dB[2];
   const double dBYdz = fl->dB[3], dBZdx = fl->dB[4], dBZdy = fl->dB[5];
   unsigned int substeps = (unsigned int)(fabs(dt) / fabs(gyro_period * (0.1 / 360.0)));
   if (substeps < 1) substeps = 1;
   const double substeps_radians = -(2.0 * VO_PI * dt / gyro_period) / substeps;
   const double substeps_dt = dt / substeps;
   double T[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
   double R[9];
   rotation_matrix(u, substeps_radians, R);
   for (unsigned int s = 0; s < substeps; ++s) {
      double p[3];
      affine_apply(T, fl->V, p);
      p[0] -= hallPrefactor * (dBZdy - dBYdz);
      p[1] -= hallPrefactor * (dBXdz - dBZdx);
      p[2] -= hallPrefactor * (dBYdx - dBXdy);
      /* Translation(-p) * T */
      for (int i = 0; i < 3; ++i) T[i * 4 + 3] = -p[i] + T[i * 4 + 3];
      /* R * T */
      double N[12];
      for (int i = 0; i < 3; ++i) {
         for (int j = 0; j < 4; ++j)
            N[i * 4 + j] = (R[i * 3 + 0] * T[0 * 4 + j] + R[i * 3 + 1] * T[1 * 4 + j]) + R[i * 3 + 2] * T[2 * 4 + j];
      }
      memcpy(T, N, sizeof(T));
      /* Translation(p) * T */
      for (int i = 0; i < 3; ++i) T[i * 4 + 3] = p[i] + T[i * 4 + 3];
      if (ohm_gradpe > 0) {
         const double qm = fabs(charge) / mass;
         for (int i = 0; i < 3; ++i) T[i * 4 + 3] = (qm * fl->EgradPe[i]) * substeps_dt + T[i * 4 + 3];
      }
   }
   memcpy(out, T, sizeof(T));
}

void vo_affine_inverse(const double in[12], double out[12]) {
   const double a = in[0], b = in[1], c = in[2];
   const double d = in[4], e = in[5], f = in[6];
   const double g = in[8], h = in[9], i = in[10];
   const double A = e * i - f * h, B = f * g - d * i, C = d * h - e * g;
   const double det = a * A + b * B + c * C;
   const double inv[9] = {
      A / det, (c * h - b * i) / det, (b * f - c * e) / det,
      B / det, (a * i - c * g) / det, (c * d - a * f) / det,
      C / det, (b * g - a * h) / det, (a * e - b * d) / det};
   for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 3; ++k) out[r * 4 + k] = inv[r * 3 + k];
      out[r * 4 + 3] = -((inv[r * 3 + 0] * in[3] + inv[r * 3 + 1] * in[7]) + inv[r * 3 + 2] * in[11]);
   }
}

/* cpu_acc_intersections.cpp:112-117 */
static void line_plane(const double lp[3], const double ld[3], const double pp[3], const double pn[3], double out[3]) {
   const double d0 = pp[0] - lp[0], d1 = pp[1] - lp[1], d2 = pp[2] - lp[2];
   const double nom = (pn[0] * d0 + pn[1] * d1) + pn[2] * d2;
   const double dem = (pn[0] * ld[0] + pn[1] * ld[1]) + pn[2] * ld[2];
   const double t = nom / dem;
   for (int q = 0; q < 3; ++q) out[q] = lp[q] + t * ld[q];
}
static void vadd(const double a[3], const double b[3], double o[3]) { o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2]; }

/* four intersections of lines/planes -> I, di, dj, dk along `dim` */
static void isect_out(const double p000[3], const double p100[3], const double p010[3], const double p001[3],
                      int dim, double* out4) {
   out4[0] = p000[dim];
   out4[1] = p100[dim] - p000[dim];
   out4[2] = p010[dim] - p000[dim];
   out4[3] = p001[dim] - p000[dim];
}

/* first intersection, cpu_acc_intersections.cpp:132-231 */
static void isect_1st(const vo_mesh* m, const double bwd[12], int dim, double* out4) {
   const double dv[3] = {vo_cell_size(m, 0), vo_cell_size(m, 1), vo_cell_size(m, 2)};
   const double* mn = m->vmin;
   double e[3] = {0, 0, 0};
   e[dim] = 1.0;
   double pn[3], pp[3], ld[3] = {0, 0, 0}, lp[3];
   linear_apply(bwd, e, pn);
   double base[3] = {0, 0, 0};
   base[dim] = mn[dim];
   affine_apply(bwd, base, pp);
   ld[dim] = 1.0;
   for (int q = 0; q < 3; ++q) lp[q] = (q == dim) ? 0.0 : 0.5 * dv[q] + mn[q];
   /* Lagrangian step along `dim`, Euclidean steps along the others */
   double step[3][3];
   for (int q = 0; q < 3; ++q) {
      double w[3] = {0, 0, 0};
      w[q] = dv[q];
      if (q == dim) linear_apply(bwd, w, step[q]);
      else { step[q][0] = w[0]; step[q][1] = w[1]; step[q][2] = w[2]; }
   }
   double r000[3], r[3][3];
   line_plane(lp, ld, pp, pn, r000);
   for (int q = 0; q < 3; ++q) {
      double lq[3], pq[3];
      if (q == dim) { lq[0] = lp[0]; lq[1] = lp[1]; lq[2] = lp[2]; vadd(pp, step[q], pq); }
      else { vadd(lp, step[q], lq); pq[0] = pp[0]; pq[1] = pp[1]; pq[2] = pp[2]; }
      line_plane(lq, ld, pq, pn, r[q]);
   }
   isect_out(r000, r[0], r[1], r[2], dim, out4);
}

/* second intersection, cpu_acc_intersections.cpp:246-358 */
static void isect_2nd(const vo_mesh* m, const double bwd[12], int dim, double* out4) {
   const double dv[3] = {vo_cell_size(m, 0), vo_cell_size(m, 1), vo_cell_size(m, 2)};
   const double* mn = m->vmin;
   /* the Euclidean plane family is normal to axis `pa`; the line follows Lagrangian axis `pa` */
   int pa;
   if (dim == 0) pa = 1;          /* order Z-X-Y: plane normal y */
   else if (dim == 1) pa = 2;     /* order X-Y-Z: plane normal z */
   else pa = 0;                   /* order Y-Z-X: plane normal x */
   double pn[3] = {0, 0, 0};
   pn[pa] = 1.0;
   double pp[3] = {0, 0, 0};
   pp[pa] = mn[pa] + dv[pa] * 0.5;
   double e[3] = {0, 0, 0};
   e[pa] = 1.0;
   double ld[3];
   linear_apply(bwd, e, ld);
   /* line point: mapped axis at the mesh minimum, the others at the first cell centre */
   double lpt[3];
   for (int q = 0; q < 3; ++q) lpt[q] = (q == dim) ? mn[q] : 0.5 * dv[q] + mn[q];
   double lp[3];
   affine_apply(bwd, lpt, lp);
   double r000[3], r[3][3];
   line_plane(lp, ld, pp, pn, r000);
   for (int q = 0; q < 3; ++q) {
      double w[3] = {0, 0, 0};
      w[q] = dv[q];
      double lq[3], pq[3];
      if (q == pa) {   /* Euclidean plane step */
         lq[0] = lp[0]; lq[1] = lp[1]; lq[2] = lp[2];
         vadd(pp, w, pq);
      } else {         /* Lagrangian line step */
         double s[3];
         linear_apply(bwd, w, s);
         vadd(lp, s, lq);
         pq[0] = pp[0]; pq[1] = pp[1]; pq[2] = pp[2];
      }
      line_plane(lq, ld, pq, pn, r[q]);
   }
   isect_out(r000, r[0], r[1], r[2], dim, out4);
}

/* third intersection, cpu_acc_intersections.cpp:376-450 */
static void isect_3rd(const vo_mesh* m, const double bwd[12], int dim, double* out4) {
   const double dv[3] = {vo_cell_size(m, 0), vo_cell_size(m, 1), vo_cell_size(m, 2)};
   const double* mn = m->vmin;
   double base[3];
   for (int q = 0; q < 3; ++q) base[q] = (q == dim ? 0.0 : 0.5) * dv[q] + mn[q];
   double pts[4][3];
   affine_apply(bwd, base, pts[0]);
   for (int q = 0; q < 3; ++q) {
      double w[3];
      for (int r = 0; r < 3; ++r) {
         double frac = (r == dim ? 0.0 : 0.5);
         if (r == q) frac += 1.0;
         w[r] = frac * dv[r] + mn[r];
      }
      affine_apply(bwd, w, pts[q + 1]);
   }
   isect_out(pts[0], pts[1], pts[2], pts[3], dim, out4);
}

void vo_intersections(const vo_mesh* m, const double fwd[12], int map_order, double out[12]) {
   static const int orders[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
   double bwd[12];
   vo_affine_inverse(fwd, bwd);
   const int* o = orders[map_order];
   isect_1st(m, bwd, o[0], out + 4 * o[0]);
   isect_2nd(m, bwd, o[1], out + 4 * o[1]);
   isect_3rd(m, bwd, o[2], out + 4 * o[2]);
}

double vo_max_v_dt(const double B[3], double mass, double charge, double max_rotation_deg) {
   const double B_mag = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]) + 1e-30;
   const double gyro_period = 2 * VO_PI * mass / (charge * B_mag);
   return fabs(gyro_period) * (max_rotation_deg / 360.0);
}

/* ------------------------------------------------------------------------ */
/* Synthetic initial condition                                               */
/* ------------------------------------------------------------------------ */

#define VO_K_B 1.3806503e-23

double vo_peak_value(const vo_peak* p, double mass, double vx, double vy, double vz) {
   const double x = vx - p->V[0], y = vy - p->V[1], z = vz - p->V[2];
   if (p->kind == 0) {
      /* projects/project.h:66-72 with Tx=Ty=Tz=T */
      const double T = p->T;
      return p->n * pow(mass / (2.0 * VO_PI * VO_K_B), 1.5) *
             exp(-mass * (x * x / T + y * y / T + z * z / T) / (2.0 * VO_K_B)) / sqrt(T * T * T);
   }
   /* gyrotropic ring about p->axis through p->V, normalised to density p->n for vth << vperp */
   const double par = x * p->axis[0] + y * p->axis[1] + z * p->axis[2];
   const double px = x - par * p->axis[0], py = y - par * p->axis[1], pz = z - par * p->axis[2];
   const double perp = sqrt(px * px + py * py + pz * pz);
   const double s2 = p->vth * p->vth;
   const double norm = 1.0 / ((2.0 * VO_PI) * (2.0 * VO_PI) * p->vperp * s2);
   const double w = perp - p->vperp;
   return p->n * norm * exp(-(w * w + par * par) / (2.0 * s2));
}

/* velocity radius beyond which a peak is below `thr` */
static double peak_radius(const vo_peak* p, double mass, double thr) {
   if (p->kind == 0) {
      const double fmax = p->n * pow(mass / (2.0 * VO_PI * VO_K_B * p->T), 1.5);
      if (fmax <= thr) return 0.0;
      return sqrt(2.0 * VO_K_B * p->T / mass * log(fmax / thr));
   }
   const double s2 = p->vth * p->vth;
   const double fmax = p->n / ((2.0 * VO_PI) * (2.0 * VO_PI) * p->vperp * s2);
   if (fmax <= thr) return 0.0;
   return p->vperp + p->vth * sqrt(2.0 * log(fmax / thr));
}

int vo_init_cell(const vo_mesh* m, const vo_peak* peaks, int npeaks, double mass, double min_value,
                 int add_width, uint32_t cap, uint32_t* n_out, uint32_t* gids_out, float* f_out) {
   /* candidate box: union of peak balls at thr/16, plus add_width+1 blocks, clamped */
   int lo[3], hi[3];
   int any = 0;
   for (int d = 0; d < 3; ++d) { lo[d] = (int)m->len[d]; hi[d] = -1; }
   const double bs[3] = {vo_cell_size(m, 0) * VO_WID, vo_cell_size(m, 1) * VO_WID, vo_cell_size(m, 2) * VO_WID};
   for (int q = 0; q < npeaks; ++q) {
      const double r = peak_radius(&peaks[q], mass, min_value / 16.0);
      if (r <= 0.0) continue;
      any = 1;
      for (int d = 0; d < 3; ++d) {
         int a = (int)floor((peaks[q].V[d] - r - m->vmin[d]) / bs[d]) - (add_width + 1);
         int b = (int)floor((peaks[q].V[d] + r - m->vmin[d]) / bs[d]) + (add_width + 1);
         if (a < lo[d]) lo[d] = a;
         if (b > hi[d]) hi[d] = b;
      }
   }
   if (!any) { *n_out = 0; return 0; }
   for (int d = 0; d < 3; ++d) {
      if (lo[d] < 0) lo[d] = 0;
      if (hi[d] > (int)m->len[d] - 1) hi[d] = (int)m->len[d] - 1;
   }
   const uint32_t nb = (uint32_t)((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1));
   uint32_t* g = (uint32_t*)malloc(sizeof(uint32_t) * nb);
   float* f = (float*)malloc(sizeof(float) * VO_WID3 * (size_t)nb);
   const double dvx = vo_cell_size(m, 0), dvy = vo_cell_size(m, 1), dvz = vo_cell_size(m, 2);
   uint32_t o = 0;
   for (int bz = lo[2]; bz <= hi[2]; ++bz)
      for (int by = lo[1]; by <= hi[1]; ++by)
         for (int bx = lo[0]; bx <= hi[0]; ++bx) {
            g[o] = (uint32_t)bx + (uint32_t)by * m->len[0] + (uint32_t)bz * m->len[0] * m->len[1];
            const double x0 = m->vmin[0] + bx * bs[0], y0 = m->vmin[1] + by * bs[1], z0 = m->vmin[2] + bz * bs[2];
            for (int k = 0; k < VO_WID; ++k)
               for (int j = 0; j < VO_WID; ++j)
                  for (int i = 0; i < VO_WID; ++i) {
                     float value = 0.0f;   /* MultiPeak.cpp:205-214, Realf accumulation */
                     for (int q = 0; q < npeaks; ++q)
                        value += (float)vo_peak_value(&peaks[q], mass, x0 + (i + 0.5) * dvx,
                                                      y0 + (j + 0.5) * dvy, z0 + (k + 0.5) * dvz);
                     f[(size_t)o * VO_WID3 + k * VO_WID2 + j * VO_WID + i] = value;
                  }
            o++;
         }
   double loss = 0.0;
   const int rc = vo_adjust(m, nb, g, f, min_value, add_width, cap, n_out, gids_out, f_out, &loss);
   free(g);
   free(f);
   return rc;
}

/* ------------------------------------------------------------------------ */
/* Batched OpenMP driver, cpu_acc_semilag.cpp:45-76                           */
/* ------------------------------------------------------------------------ */

int64_t vo_accelerate_cells(const vo_mesh* m, int ncells, const uint32_t* offsets,
                            const uint32_t* gids, const float* f, const double* intersections,
                            int map_order, int method, double min_value, int wall_margin,
                            int do_adjust, int add_width, int nthreads,
                            uint64_t cap, uint32_t* out_offsets, uint32_t* gids_out, float* f_out,
                            int* wall_flags) {
   vo_cell* cells = (vo_cell*)calloc((size_t)ncells, sizeof(vo_cell));
#ifdef _OPENMP
   if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
   for (int c = 0; c < ncells; ++c) {
      const uint32_t n = offsets[c + 1] - offsets[c];
      vo_cell_init(&cells[c], n, gids + offsets[c], f + (size_t)offsets[c] * VO_WID3);
      int wf = 0;
      vo_accelerate_cell(m, &cells[c], intersections + 12 * (size_t)c, map_order, method, min_value,
                         wall_margin, &wf);
      if (do_adjust) {
         uint32_t cap2 = cells[c].n * 2 + 1024, nn = 0;
         uint32_t* g2 = (uint32_t*)malloc(sizeof(uint32_t) * cap2);
         float* f2 = (float*)malloc(sizeof(float) * VO_WID3 * (size_t)cap2);
         double loss;
         while (vo_adjust(m, cells[c].n, cells[c].gids, cells[c].f, min_value, add_width, cap2, &nn, g2, f2, &loss) != 0) {
            free(g2); free(f2);
            cap2 = nn + 16;
            g2 = (uint32_t*)malloc(sizeof(uint32_t) * cap2);
            f2 = (float*)malloc(sizeof(float) * VO_WID3 * (size_t)cap2);
         }
         free(cells[c].gids); free(cells[c].f);
         cells[c].gids = g2; cells[c].f = f2; cells[c].n = nn; cells[c].cap = cap2;
      }
      if (wall_flags) wall_flags[c] = wf;
   }
   uint64_t tot = 0;
   for (int c = 0; c < ncells; ++c) {
      if (out_offsets) out_offsets[c] = (uint32_t)tot;
      tot += cells[c].n;
   }
   if (out_offsets) out_offsets[ncells] = (uint32_t)tot;
   int64_t ret = (int64_t)tot;
   if (gids_out && f_out && cap > 0) {
      if (tot > cap) ret = -1;
      else {
         for (int c = 0; c < ncells; ++c) {
            memcpy(gids_out + out_offsets[c], cells[c].gids, sizeof(uint32_t) * cells[c].n);
            memcpy(f_out + (size_t)out_offsets[c] * VO_WID3, cells[c].f, sizeof(float) * VO_WID3 * cells[c].n);
         }
      }
   }
   for (int c = 0; c < ncells; ++c) vo_cell_free(&cells[c]);
   free(cells);
   return ret;
}

#ifndef VO_REALF_DOUBLE
/* ------------------------------------------------------------------------ */
/* 16-bit block storage: codes plus a per-block FP32 scale (the library's     */
/* BF16 / FP16 storage forms).  No reference counterpart: the reference       */
/* stores Realf.  Restated from the format definition:                        */
/*   m = max|x| = mant*2^e (mant in [0.5,1)); code = RNE(x * 2^-e);           */
/*   S_x, S_d = canonical FP32 sums of x and of the decoded codes;             */
/*   scale = S_x/S_d if S_x > 0, S_d > 0 and 2^(e-1) <= S_x/S_d <= 2^(e+1),    */
/*   else 2^e (0 for an all-zero block); f = scale * code.                     */
/* ------------------------------------------------------------------------ */

static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

uint16_t vo_f32_to_bf16(float x) {
   const uint32_t u = f2u(x);
   return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);   /* round to nearest even, finite x */
}

float vo_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

/* IEEE binary32 -> binary16, round to nearest even, with subnormal results (finite x). */
uint16_t vo_f32_to_f16(float x) {
   const uint32_t u = f2u(x);
   const uint32_t sign = (u >> 16) & 0x8000u;
   const uint32_t a = u & 0x7FFFFFFFu;
   if (a >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);     /* >= 65520: overflow to inf */
   if (a >= 0x38800000u) {                                          /* normal binary16 */
      uint32_t h = (a - 0x38000000u) >> 13;
      const uint32_t rem = a & 0x1FFFu;
      if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
      return (uint16_t)(sign | h);
   }
   if (a < 0x33000000u) return (uint16_t)sign;                     /* < 2^-25: rounds to zero */
   /* subnormal binary16: value / 2^-24 rounded to an integer */
   const uint32_t e = a >> 23;
   const uint32_t mant = (a & 0x7FFFFFu) | 0x800000u;               /* value = mant * 2^(e-150) */
   const uint32_t shift = 126u - e;                                 /* 14..24 */
   uint32_t q = mant >> shift;
   const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
   if (rem > half || (rem == half && (q & 1u))) ++q;
   return (uint16_t)(sign | q);
}

float vo_f16_to_f32(uint16_t h) {
   const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
   const uint32_t e = ((uint32_t)h >> 10) & 0x1Fu, mant = (uint32_t)h & 0x3FFu;
   if (e == 0) {
      const float v = (float)mant * u2f(0x33800000u);               /* mant * 2^-24, exact */
      return sign ? -v : v;
   }
   if (e == 31) return u2f(sign | 0x7F800000u | (mant << 13));
   return u2f(sign | ((e + 112u) << 23) | (mant << 13));
}

static uint16_t enc16(int st, float x) { return st == VO_ST_BF16 ? vo_f32_to_bf16(x) : vo_f32_to_f16(x); }
static float dec16(int st, uint16_t h) { return st == VO_ST_BF16 ? vo_bf16_to_f32(h) : vo_f16_to_f32(h); }

/* canonical sum of a block: rows L of (i = 0..3) then the 16 row sums pairwise s_L + s_{L^m},
 * m = 8, 4, 2, 1 */
float vo_block_csum(const float x[64]) {
   float s[16], t[16];
   for (int L = 0; L < 16; ++L) s[L] = (x[4 * L] + x[4 * L + 1]) + (x[4 * L + 2] + x[4 * L + 3]);
   for (int m = 8; m > 0; m >>= 1) {
      for (int L = 0; L < 16; ++L) t[L] = s[L] + s[L ^ m];
      memcpy(s, t, sizeof(s));
   }
   return s[0];
}

void vo_encode_block(const float x[64], int storage, uint16_t code[64], float* scale) {
   float m = 0.0f;
   for (int e = 0; e < 64; ++e) m = fmaxf(m, fabsf(x[e]));
   if (!(m > 0.0f)) {
      memset(code, 0, 64 * sizeof(uint16_t));
      *scale = 0.0f;
      return;
   }
   int ex;
   (void)frexpf(m, &ex);                     /* m = mant * 2^ex, mant in [0.5, 1) */
   const float p = ldexpf(1.0f, ex);
   float d[64];
   for (int e = 0; e < 64; ++e) {
      code[e] = enc16(storage, ldexpf(x[e], -ex));   /* exact power-of-two normalisation */
      d[e] = dec16(storage, code[e]);
   }
   const float sf = vo_block_csum(x), sq = vo_block_csum(d);
   const float r = (sq > 0.0f) ? sf / sq : 0.0f;
   *scale = (sf > 0.0f && r >= 0.5f * p && r <= 2.0f * p) ? r : p;
}

void vo_decode_block(const uint16_t code[64], float scale, int storage, float out[64]) {
   for (int e = 0; e < 64; ++e) out[e] = scale * dec16(storage, code[e]);
}

void vo_convert16(int storage, const float* x, uint32_t n, uint16_t* out) {
   for (uint32_t i = 0; i < n; ++i) out[i] = enc16(storage, x[i]);
}

void vo_requantize(uint32_t nblocks, float* f, int storage) {
   uint16_t code[64];
   float sc;
   for (uint32_t b = 0; b < nblocks; ++b) {
      vo_encode_block(f + (size_t)b * 64, storage, code, &sc);
      vo_decode_block(code, sc, storage, f + (size_t)b * 64);
   }
}
#endif /* !VO_REALF_DOUBLE */

/* ------------------------------------------------------------------------ */
/* Spatial translation (cpu_trans_map_amr.cpp, cpu_1d_ppm_nonuniform.hpp)    */
/* ------------------------------------------------------------------------ */

/* h4 left face value on non-uniform cell widths h, cpu_face_estimates.hpp:179-191
 * (White et al. 2008 eq. 45).  The h-only sub-expressions are float arithmetic except the
 * terms that involve 1.0 or 2.0, which are double; each double scalar multiplies the Vec in
 * double and rounds into the lane (dm). */
static float h4nu_face(const float* h, const float* u, int k) {
   const float s4 = ((h[k - 2] + h[k - 1]) + h[k]) + h[k + 1];
   const double D1 = 1.0 / (double)s4;
   const float SA = ((h[k - 2] + h[k - 1]) * (h[k] + h[k + 1])) / (h[k - 1] + h[k]);
   const float V1 = u[k - 1] * h[k] + u[k] * h[k - 1];
   const double DA = 1.0 / (double)((h[k - 2] + h[k - 1]) + h[k]) + 1.0 / (double)((h[k - 1] + h[k]) + h[k + 1]);
   const float A = dm(DA, SA * V1);
   const float SB = (h[k] * (h[k] + h[k + 1])) / (((h[k - 2] + h[k - 1]) + h[k]) * (h[k - 2] + h[k - 1]));
   const double DB = (double)h[k - 2] + 2.0 * (double)h[k - 1];
   const float V2 = dm(DB, u[k - 1]) - u[k - 2] * h[k - 1];
   const float B = SB * V2;
   const float SC = (h[k - 1] * (h[k - 2] + h[k - 1])) / (((h[k - 1] + h[k]) + h[k + 1]) * (h[k] + h[k + 1]));
   const double DC = 2.0 * (double)h[k] + (double)h[k + 1];
   const float V3 = dm(DC, u[k]) - u[k + 1] * h[k];
   const float C = SC * V3;
   return dm(D1, (A + B) + C);
}

/* Filtered faces on non-uniform widths, cpu_face_estimates.hpp:327-372 (order h4) */
static void ppm_nu_faces(const float* h, const float* v, int k, float thr, float* pm, float* pp) {
   float fvl = h4nu_face(h, v, k);
   float fvr = h4nu_face(h, v, k + 1);
   float sabs, ssign;
   if (thr > 0.0f) {
      limiter_scaled(v, k, thr, &sabs, &ssign);
   } else {
      const float s = superbee(v[k - 1], v[k], v[k + 1]);
      sabs = fabsf(s);
      ssign = (s > 0.0f) ? 1.0f : -1.0f;
   }
   if (sabs == 0.0f) {
      fvr = v[k];
      fvl = v[k];
   }
   {
      const float p = (v[k - 1] - fvl) * (fvl - v[k]);
      if (p < 0.0f) fvl = v[k] - dm(0.5, ssign) * sabs;
   }
   {
      const float p = (v[k + 1] - fvr) * (fvr - v[k]);
      if (p < 0.0f) fvr = v[k] + dm(0.5, ssign) * sabs;
   }
   *pm = fvl;
   *pp = fvr;
}

/* PPM coefficients on non-uniform widths, cpu_1d_ppm_nonuniform.hpp:40-62 */
void vo_ppm_nu_coeff(const float* h, const float* v, int k, float thr, float a[3]) {
   float m, p;
   ppm_nu_faces(h, v, k, thr, &m, &p);
   const float one_sixth = (float)(1.0 / 6.0);
   {
      const float d = p - m;
      const float mid = v[k] - dm(0.5, m + p);
      if (d * mid > (d * d) * one_sixth) m = 3.0f * v[k] - 2.0f * p;
   }
   {
      const float d = p - m;
      const float mid = v[k] - dm(0.5, m + p);
      if (((-d) * d) * one_sixth > d * mid) p = 3.0f * v[k] - 2.0f * m;
   }
   a[0] = m;
   a[1] = (dm(3.0, v[k]) - dm(2.0, m)) - p;
   a[2] = (m + p) - dm(2.0, v[k]);
}

void vo_ppm_nu_batch(const float* h5, const float* stencils, int n, float thr, float* out) {
   for (int i = 0; i < n; ++i) vo_ppm_nu_coeff(h5, stencils + 5 * i, 2, thr, out + 3 * i);
}

typedef struct { uint32_t gid, idx; } vo_gi;

static int cmp_gi(const void* a, const void* b) {
   const uint32_t x = ((const vo_gi*)a)->gid, y = ((const vo_gi*)b)->gid;
   return (x > y) - (x < y);
}

/* Index of block gid in cell c (sorted table), or VO_INVALID. */
static uint32_t find_block(const vo_gi* tab, const uint64_t* toff, uint32_t c, uint32_t gid) {
   uint64_t lo = toff[c], hi = toff[c + 1];
   while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if (tab[mid].gid < gid) lo = mid + 1;
      else hi = mid;
   }
   return (lo < toff[c + 1] && tab[lo].gid == gid) ? tab[lo].idx : VO_INVALID;
}

/* Translation of every cell along dimension dim by dt over the given pencils
 * (trans_map_1d_amr, cpu_trans_map_amr.cpp:214-441, with propagatePencil :36-149 and
 * copy_trans_block_data_amr :162-198).  Pencil p covers positions [pstart[p],
 * pstart[p] + plen[p]) of pcell (cell index or VO_INVALID), pdz (sourceDZ) and pratio
 * (targetRatios); its first and last VO_TRANS_STENCIL positions are stencil cells.
 * Pencils come in the reference's processing order (bins, then pencils in a bin).
 * The block lists do not change; f is replaced in place.  Cells that are a target
 * (ratio > 0) at any pencil position start from zero and receive the contributions in
 * pencil order, then source position order; other cells keep their values. */
int vo_trans_map_1d(const vo_mesh* m, uint32_t ncells, const uint64_t* offsets, const uint32_t* gids, float* f,
                    uint32_t npencils, const uint32_t* plen, const uint64_t* pstart, const uint32_t* pcell,
                    const float* pdz, const float* pratio, int dim, float dt, float threshold) {
   const int W = VO_TRANS_STENCIL;
   const uint64_t nb = offsets[ncells];
   /* per-cell sorted GID tables */
   vo_gi* tab = (vo_gi*)malloc((nb ? nb : 1) * sizeof(vo_gi));
   float* src = (float*)malloc((nb ? nb : 1) * VO_WID3 * sizeof(float));
   uint8_t* is_target = (uint8_t*)calloc(ncells ? ncells : 1, 1);
   if (!tab || !src || !is_target) {
      free(tab);
      free(src);
      free(is_target);
      return -1;
   }
   for (uint32_t c = 0; c < ncells; ++c) {
      for (uint64_t b = offsets[c]; b < offsets[c + 1]; ++b) {
         tab[b].gid = gids[b];
         tab[b].idx = (uint32_t)b;
      }
      qsort(tab + offsets[c], offsets[c + 1] - offsets[c], sizeof(vo_gi), cmp_gi);
   }
   memcpy(src, f, nb * VO_WID3 * sizeof(float));
   /* targetCellsInBin (cpu_trans_pencils.hpp:49, binPencils): cells written by any pencil */
   for (uint32_t p = 0; p < npencils; ++p)
      for (uint32_t q = 0; q < plen[p]; ++q) {
         const uint32_t c = pcell[pstart[p] + q];
         if (c != VO_INVALID && pratio[pstart[p] + q] > 0.0f) is_target[c] = 1;
      }
   for (uint32_t c = 0; c < ncells; ++c)
      if (is_target[c]) memset(f + offsets[c] * VO_WID3, 0, (offsets[c + 1] - offsets[c]) * VO_WID3 * sizeof(float));
   /* union of block GIDs over all cells */
   uint32_t* uni = (uint32_t*)malloc((nb ? nb : 1) * sizeof(uint32_t));
   for (uint64_t b = 0; b < nb; ++b) uni[b] = gids[b];
   qsort(uni, nb, sizeof(uint32_t), cmp_u32);
   uint64_t nu = 0;
   for (uint64_t b = 0; b < nb; ++b)
      if (nu == 0 || uni[nu - 1] != uni[b]) uni[nu++] = uni[b];
   /* vcell_transpose, cpu_trans_map_amr.cpp:224-285 */
   uint32_t to_id[3];
   if (dim == 0) { to_id[0] = VO_WID2; to_id[1] = VO_WID; to_id[2] = 1; }
   else if (dim == 1) { to_id[0] = 1; to_id[1] = VO_WID2; to_id[2] = VO_WID; }
   else { to_id[0] = 1; to_id[1] = VO_WID; to_id[2] = VO_WID2; }
   uint32_t vt[VO_WID3];
   for (uint32_t k = 0; k < VO_WID; ++k)
      for (uint32_t j = 0; j < VO_WID; ++j)
         for (uint32_t i = 0; i < VO_WID; ++i) vt[i + j * VO_WID + k * VO_WID2] = i * to_id[0] + j * to_id[1] + k * to_id[2];
   const float dvz = (float)vo_cell_size(m, dim);
   const float vz_min = (float)m->vmin[dim];
   uint32_t maxlen = 0;
   for (uint32_t p = 0; p < npencils; ++p) maxlen = plen[p] > maxlen ? plen[p] : maxlen;
   float* vals = (float*)malloc(((size_t)maxlen + 1) * VO_WID3 * sizeof(float));   /* [position][transposed cell] */
   uint32_t* bidx = (uint32_t*)malloc(((size_t)maxlen + 1) * sizeof(uint32_t));
   for (uint32_t p = 0; p < npencils; ++p) {
      const uint32_t L = plen[p];
      const uint64_t s0 = pstart[p];
      const float* dz = pdz + s0;
      const float* ratio = pratio + s0;
      for (uint64_t gi = 0; gi < nu; ++gi) {
         const uint32_t g = uni[gi];
         int nonempty = 0;
         for (uint32_t b = 0; b < L; ++b) {
            const uint32_t c = pcell[s0 + b];
            bidx[b] = (c == VO_INVALID) ? VO_INVALID : find_block(tab, offsets, c, g);
            if (bidx[b] != VO_INVALID) {
               ++nonempty;
               for (int t = 0; t < VO_WID3; ++t) vals[b * VO_WID3 + t] = src[(uint64_t)bidx[b] * VO_WID3 + vt[t]];
            } else {
               for (int t = 0; t < VO_WID3; ++t) vals[b * VO_WID3 + t] = 0.0f;
            }
         }
         if (!nonempty) continue;
         const uint32_t bz = (dim == 0) ? g % m->len[0] : (dim == 1) ? (g / m->len[0]) % m->len[1] : g / (m->len[0] * m->len[1]);
         for (int i = W; i < (int)L - W; ++i) {
            const uint32_t tb[3] = {bidx[i - 1], bidx[i], bidx[i + 1]};
            for (uint32_t k = 0; k < VO_WID; ++k) {
               const float cell_vz = (float)(((double)(bz * VO_WID + k) + 0.5) * (double)dvz + (double)vz_min);
               const float zt = (cell_vz * dt) / dz[i];
               const int pos = zt > 0.0f;
               const float z1 = pos ? (float)(1.0 - (double)zt) : 0.0f;
               const float z2 = pos ? 1.0f : -zt;
               for (uint32_t pv = 0; pv < VO_WID2 / 8; ++pv) {
                  /* check_skip_remapping, cpu_trans_map_amr.cpp:18-25: 8 lanes x stencil */
                  int any = 0;
                  for (int s = -W; s <= W && !any; ++s)
                     for (int l = 0; l < 8; ++l)
                        if (vals[(i + s) * VO_WID3 + k * VO_WID2 + pv * 8 + l] > 0.0f) any = 1;
                  if (!any) continue;
                  for (int l = 0; l < 8; ++l) {
                     const int t = k * VO_WID2 + pv * 8 + l;
                     float u[5], h[5], a[3];
                     for (int s = 0; s < 5; ++s) {
                        u[s] = vals[(i - 2 + s) * VO_WID3 + t];
                        h[s] = dz[i - 2 + s];
                     }
                     vo_ppm_nu_coeff(h, u, 2, threshold, a);
                     const float ng = z2 * (a[0] + z2 * (a[1] + z2 * a[2])) - z1 * (a[0] + z1 * (a[1] + z1 * a[2]));
                     if (ratio[i] != 0.0f && tb[1] != VO_INVALID)
                        f[(uint64_t)tb[1] * VO_WID3 + vt[t]] += (u[2] - ng) * ratio[i];
                     if (ratio[i + 1] != 0.0f && tb[2] != VO_INVALID)
                        f[(uint64_t)tb[2] * VO_WID3 + vt[t]] += (pos ? (ng * dz[i]) / dz[i + 1] : 0.0f) * ratio[i + 1];
                     if (ratio[i - 1] != 0.0f && tb[0] != VO_INVALID)
                        f[(uint64_t)tb[0] * VO_WID3 + vt[t]] += (!pos ? (ng * dz[i]) / dz[i - 1] : 0.0f) * ratio[i - 1];
                  }
               }
            }
         }
      }
   }
   free(vals);
   free(bidx);
   free(uni);
   free(tab);
   free(src);
   free(is_target);
   return 0;
}
