/* CPU oracle for the cosmic-collisions simulation step -- TEST INFRASTRUCTURE.
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg load this
 * library, and only as the checker; the product never does.  Every function
 * restates one reference pass in plain C: FP32 where the reference is a GLSL
 * shader (operation order kept as written in the shader, no contraction:
 * build with -ffp-contract=off), FP64 where the reference computes on the
 * host in JavaScript numbers.  GLSL built-ins are taken as:
 *   inversesqrt(x) = 1/sqrt(x) (both correctly rounded), mix(x,y,a) = x*(1-a)+y*a,
 *   smoothstep(e0,e1,x): t = clamp((x-e0)/(e1-e0),0,1), t*t*(3-2t),
 *   step(e,x) = x < e ? 0 : 1, mod(x,y) = x - y*floor(x/y).
 * How closely each function is pinned (golden vectors from the reference's
 * own passes, FP64 brute force, invariants) is stated in the test files.
 *
 * CCO_MUT (0 unless set at build time: `make mutants`) plants one known
 * mistake in the far-field and per-cell passes, so that the tests pinning
 * them can show each planted mistake fails them (tests/test_cell_physics.py).
 * The library the tests check against is built without it.
 */
#ifndef CCO_MUT
#define CCO_MUT 0
#endif
#include "cc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EMPTY_SEAT 0xBF800000u /* -1.0f: the only seat pattern with the sign bit */

static inline float rsq(float x) { return 1.0f / sqrtf(x); }
static inline float fmaxf_(float a, float b) { return a > b ? a : b; }   /* GLSL max */
static inline float fminf_(float a, float b) { return a < b ? a : b; }
static inline float clampf(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }
static inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static inline float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int shl(int v, int s) { return (int)((uint32_t)v << s); }   /* GLSL << on negative ints */

/* 3-D mesh -> 2-D atlas (src/glsl-sim.js:12-15) */
static inline int fine_at(int cx, int cy, int cz) { return ((cz >> 3) * 64 + cy) * 512 + ((cz & 7) * 64 + cx); }
static inline int coarse_at(int cx, int cy, int cz) { return ((cz >> 2) * 16 + cy) * 64 + ((cz & 3) * 16 + cx); }

void cco_grid_make(int G, cco_grid* g) {   /* src/gl.js:73-83 */
  g->G = G;
  g->GShift = G == 128 ? 7 : G == 64 ? 6 : (int)lround(log2((double)G));
  g->GMask = G - 1;
  g->SX = G == 128 ? 16 : 8;
  g->SXShift = g->SX == 16 ? 4 : 3;
  g->SXMask = g->SX - 1;
  g->AW = g->SX * G;
  g->AH = (G / g->SX) * G;
  g->IB = 18;
}

void cco_hash_at(const cco_grid* g, int cx, int cy, int cz, int* ax, int* ay) {
  int px = cx >> g->GShift, py = cy >> g->GShift, pz = cz >> g->GShift;   /* arithmetic shift */
  cy += shl(px, g->GShift - 1) + shl(pz, g->GShift - 2);
  cz += shl(py, g->GShift - 1);
  cx &= g->GMask; cy &= g->GMask; cz &= g->GMask;
  *ax = shl(cz & g->SXMask, g->GShift) + cx;
  *ay = shl(cz >> g->SXShift, g->GShift) + cy;
}

static inline uint32_t seat_word(uint32_t me, const float* gq, int ib) {
  /* uvec3(clamp(fract(g) * 15.0, 0.0, 14.0)) (src/glsl-sim.js:143-144) */
  uint32_t q[3];
  for (int a = 0; a < 3; a++) {
    float fr = gq[a] - floorf(gq[a]);
    q[a] = (uint32_t)clampf(fr * 15.0f, 0.0f, 14.0f);
  }
  /* 19-bit indices (no reference counterpart) move the place fields up one bit and drop bit 30 */
  return me | (q[0] << ib) | (q[1] << (ib + 4)) | (q[2] << (ib + 8)) | (ib == 18 ? 0x40000000u : 0u);
}

/* The depth-peeled seat lists: seat k of a hashed cell holds the k-th lowest
 * particle index whose cell hashes there (pass k keeps, under a LESS depth
 * test at depth i/N, the lowest index above seat k-1's).  Walking the
 * particles in index order and filling each cell's seats first-come is the
 * same set in the same order. */
void cco_slots(const float* pos, int N, float invCell, const cco_grid* g, uint32_t* seatA, uint32_t* seatB) {
  size_t nt = (size_t)g->AW * g->AH;
  for (size_t t = 0; t < nt * 4; t++) { seatA[t] = EMPTY_SEAT; seatB[t] = EMPTY_SEAT; }
  unsigned char* cnt = (unsigned char*)calloc(nt, 1);
  for (int i = 0; i < N; i++) {
    float gq[3] = { pos[i * 4] * invCell, pos[i * 4 + 1] * invCell, pos[i * 4 + 2] * invCell };
    int ax, ay;
    cco_hash_at(g, (int)floorf(gq[0]), (int)floorf(gq[1]), (int)floorf(gq[2]), &ax, &ay);
    size_t t = (size_t)ay * g->AW + ax;
    int k = cnt[t];
    if (k >= 8) continue;
    cnt[t] = (unsigned char)(k + 1);
    uint32_t w = seat_word((uint32_t)i, gq, g->IB);
    if (k & 1) seatB[t * 4 + (k >> 1)] = w; else seatA[t * 4 + (k >> 1)] = w;
  }
  free(cnt);
}

int cco_box_of(const cco_box* b, const float p[3], float g[3]) {
  float gA[3], gB[3];
  int inA = 1, inB = b->boxB > 0.5f;
  for (int a = 0; a < 3; a++) {
    gA[a] = (p[a] + b->LA[a]) * b->invH - 0.5f;
    gB[a] = (p[a] + b->LB[a]) * b->invH - 0.5f;
    if (!(gA[a] >= 0.0f && gA[a] <= 63.0f)) inA = 0;
  }
  if (inB) for (int a = 0; a < 3; a++) if (!(gB[a] >= 0.0f && gB[a] <= 63.0f)) inB = 0;
  if (inA && inB) {
    float mA = 0.0f, mB = 0.0f;
    for (int a = 0; a < 3; a++) { mA = fmaxf_(mA, fabsf(gA[a] - 31.5f)); mB = fmaxf_(mB, fabsf(gB[a] - 31.5f)); }
    /* max(d.x, max(d.y, d.z)) */
    float dA[3], dB[3];
    for (int a = 0; a < 3; a++) { dA[a] = fabsf(gA[a] - 31.5f); dB[a] = fabsf(gB[a] - 31.5f); }
    mA = fmaxf_(dA[0], fmaxf_(dA[1], dA[2]));
    mB = fmaxf_(dB[0], fmaxf_(dB[1], dB[2]));
    inA = mA <= mB;
    inB = !inA;
  }
  for (int a = 0; a < 3; a++) g[a] = inA ? gA[a] : gB[a];
  return inA ? 0 : inB ? 1 : -1;
}

void cco_deposit(const float* pos, const float* mat, int N, const cco_box* b, int box, const float mono[4],
                 float eps2, float* fine) {
  memset(fine, 0, sizeof(float) * 513 * 512 * 4);
  float* scratch = fine + (size_t)512 * 512 * 4;   /* texel (0, 512) */
  uint32_t nv = (uint32_t)N * 8u;
  for (uint32_t v = 0; v < nv; v++) {
    int i = (int)(((v >> 3) * 40503u) & (uint32_t)(N - 1)), k = (int)(v & 7u);
    const float* p = pos + (size_t)i * 4;
    float mr = mat[i], g[3];
    int mine = cco_box_of(b, p, g);
    if (mine != box) {
      if (k != 0 || mine >= 0) continue;
      float r[3];
      for (int a = 0; a < 3; a++) r[a] = mono[a] / fmaxf_(mono[3], 1e-6f) - p[a];
      float inv = rsq(dot3(r, r) + eps2);
      float s = mr * inv * inv * inv;
      for (int a = 0; a < 3; a++) scratch[a] += r[a] * s;
      scratch[3] += mr;
      continue;
    }
    int o[3] = { k & 1, (k >> 1) & 1, k >> 2 }, c[3];
    float w3[3];
    for (int a = 0; a < 3; a++) {
      float f = floorf(g[a]), t = g[a] - f;
      c[a] = (int)f + o[a];
      w3[a] = mixf(1.0f - t, t, (float)o[a]);
    }
    if (c[0] > 63 || c[1] > 63 || c[2] > 63) continue;
    float w = w3[0] * w3[1] * w3[2] * mr;
    float* dst = fine + (size_t)fine_at(c[0], c[1], c[2]) * 4;
    dst[0] += p[0] * w; dst[1] += p[1] * w; dst[2] += p[2] * w; dst[3] += 1.0f * w;
  }
}

void cco_coarse(const float* fine, float* coarse) {
  for (int my = 0; my < 64; my++) for (int mx = 0; mx < 64; mx++) {
    int c[3] = { mx & 15, my & 15, (my >> 4) * 4 + (mx >> 4) };
    float s[4] = { 0, 0, 0, 0 };
    for (int z = 0; z < 4; z++) for (int y = 0; y < 4; y++) for (int x = 0; x < 4; x++) {
      const float* f = fine + (size_t)fine_at(c[0] * 4 + x, c[1] * 4 + y, c[2] * 4 + z) * 4;
      for (int a = 0; a < 4; a++) s[a] += f[a];
    }
    memcpy(coarse + (size_t)(my * 64 + mx) * 4, s, sizeof s);
  }
}

void cco_mono(const float* coarse, float* mono) {
  float s[4] = { 0, 0, 0, 0 };
  for (int y = 0; y < 64; y++) for (int x = 0; x < 64; x++)
    for (int a = 0; a < 4; a++) s[a] += coarse[(size_t)(y * 64 + x) * 4 + a];
  memcpy(mono, s, sizeof s);
}

void cco_far(const float* coarse, float* farA, float* farT0, float* farT1) {
  for (int my = 0; my < 64; my++) for (int mx = 0; mx < 64; mx++) {
    size_t me = (size_t)(my * 64 + mx) * 4;
    int c[3] = { mx & 15, my & 15, (my >> 4) * 4 + (mx >> 4) };
    const float* mc = coarse + me;
    for (int a = 0; a < 4; a++) { farA[me + a] = 0; farT0[me + a] = 0; farT1[me + a] = 0; }
    if (mc[3] <= 0.0f) continue;
    float x[3] = { mc[0] / mc[3], mc[1] / mc[3], mc[2] / mc[3] };
    float acc[3] = { 0, 0, 0 }, td[3] = { 0, 0, 0 }, to[3] = { 0, 0, 0 }, pot = 0.0f;
    for (int kz = 0; kz < 16; kz++) for (int ky = 0; ky < 16; ky++) for (int kx = 0; kx < 16; kx++) {
      int dx = abs(kx - c[0]), dy = abs(ky - c[1]), dz = abs(kz - c[2]);
      int dm = dx > dy ? dx : dy; dm = dm > dz ? dm : dz;
#if CCO_MUT == 7
      if (dm <= 1) continue;
#else
      if (dm <= 2) continue;
#endif
      const float* m = coarse + (size_t)coarse_at(kx, ky, kz) * 4;
      if (m[3] <= 0.0f) continue;
#if CCO_MUT == 8
      float r[3] = { (kx * 4 + 2) * 0.1875f - 6.0f - x[0], (ky * 4 + 2) * 0.1875f - 6.0f - x[1],
                     (kz * 4 + 2) * 0.1875f - 6.0f - x[2] };
#else
      float r[3] = { m[0] / m[3] - x[0], m[1] / m[3] - x[1], m[2] / m[3] - x[2] };
#endif
      float inv2 = 1.0f / dot3(r, r), inv = sqrtf(inv2), inv3 = inv * inv2, inv5 = inv3 * inv2;
      float s = m[3] * inv3;
      for (int a = 0; a < 3; a++) acc[a] += r[a] * s;
      pot -= m[3] * inv;
      for (int a = 0; a < 3; a++) td[a] += m[3] * (3.0f * r[a] * r[a] * inv5 - inv3);
      float oxy = r[0] * r[1], oxz = r[0] * r[2], oyz = r[1] * r[2];
      to[0] += m[3] * (3.0f * oxy * inv5);
      to[1] += m[3] * (3.0f * oxz * inv5);
      to[2] += m[3] * (3.0f * oyz * inv5);
    }
    farA[me] = acc[0]; farA[me + 1] = acc[1]; farA[me + 2] = acc[2]; farA[me + 3] = pot;
    farT0[me] = td[0]; farT0[me + 1] = td[1]; farT0[me + 2] = td[2]; farT0[me + 3] = to[0];
    farT1[me] = to[1]; farT1[me + 1] = to[2];
  }
}

void cco_cell(const float* fine, const float* coarse, const float* farA, const float* farT0, const float* farT1,
              float eps2, float gm, float* force) {
  for (int my = 0; my < 513; my++) for (int mx = 0; mx < 512; mx++) {
    size_t me = ((size_t)my * 512 + mx) * 4;
    if (my >= 512) { for (int a = 0; a < 4; a++) force[me + a] = fine[me + a]; continue; }
    int c[3] = { mx & 63, my & 63, (my >> 6) * 8 + (mx >> 6) };
    const float* mc = fine + me;
    if (mc[3] <= 0.0f) { for (int a = 0; a < 4; a++) force[me + a] = 0.0f; continue; }
    float x[3] = { mc[0] / mc[3], mc[1] / mc[3], mc[2] / mc[3] };
    int cc[3] = { c[0] >> 2, c[1] >> 2, c[2] >> 2 };
    float acc[3] = { 0, 0, 0 }, pot = 0.0f;
#if CCO_MUT == 3
    eps2 = eps2 * 2.0f;
#endif
#if CCO_MUT == 4
    for (int kz = -1; kz <= 1; kz++) for (int ky = -1; ky <= 1; ky++) for (int kx = -1; kx <= 1; kx++) {
#else
    for (int kz = -2; kz <= 2; kz++) for (int ky = -2; ky <= 2; ky++) for (int kx = -2; kx <= 2; kx++) {
#endif
      int k[3] = { cc[0] + kx, cc[1] + ky, cc[2] + kz };
      if (k[0] < 0 || k[1] < 0 || k[2] < 0 || k[0] > 15 || k[1] > 15 || k[2] > 15) continue;
      if (coarse[(size_t)coarse_at(k[0], k[1], k[2]) * 4 + 3] <= 0.0f) continue;
      for (int fz = 0; fz < 4; fz++) for (int fy = 0; fy < 4; fy++) for (int fx = 0; fx < 4; fx++) {
        int f[3] = { k[0] * 4 + fx, k[1] * 4 + fy, k[2] * 4 + fz };
#if CCO_MUT != 6
        if (f[0] == c[0] && f[1] == c[1] && f[2] == c[2]) continue;
#endif
        const float* mf = fine + (size_t)fine_at(f[0], f[1], f[2]) * 4;
        if (mf[3] <= 0.0f) continue;
        float r[3] = { mf[0] / mf[3] - x[0], mf[1] / mf[3] - x[1], mf[2] / mf[3] - x[2] };
        float inv = rsq(dot3(r, r) + eps2);
        float s = mf[3] * inv * inv * inv;
        for (int a = 0; a < 3; a++) acc[a] += r[a] * s;
        pot -= mf[3] * inv;
      }
    }
    size_t ca = (size_t)coarse_at(cc[0], cc[1], cc[2]) * 4;
    const float *mb = coarse + ca, *fa = farA + ca, *t0 = farT0 + ca, *t1 = farT1 + ca;
    float d[3] = { x[0] - mb[0] / mb[3], x[1] - mb[1] / mb[3], x[2] - mb[2] / mb[3] };
#if CCO_MUT == 5
    float Td[3] = { t0[0] * d[0] + t0[3] * d[1] + t1[1] * d[2],
                    t0[3] * d[0] + t0[1] * d[1] + t1[1] * d[2],
                    t1[0] * d[0] + t1[1] * d[1] + t0[2] * d[2] };
#else
    float Td[3] = { t0[0] * d[0] + t0[3] * d[1] + t1[0] * d[2],
                    t0[3] * d[0] + t0[1] * d[1] + t1[1] * d[2],
                    t1[0] * d[0] + t1[1] * d[1] + t0[2] * d[2] };
#endif
#if CCO_MUT == 1
    for (int a = 0; a < 3; a++) acc[a] += fa[a];
#elif CCO_MUT == 2
    for (int a = 0; a < 3; a++) acc[a] += fa[a] - Td[a];
#else
    for (int a = 0; a < 3; a++) acc[a] += fa[a] + Td[a];
#endif
#if CCO_MUT == 9
    pot += fa[3] - dot3(fa, d);
#else
    pot += fa[3] - dot3(fa, d) - 0.5f * dot3(d, Td);
#endif
    force[me] = acc[0] * gm; force[me + 1] = acc[1] * gm; force[me + 2] = acc[2] * gm; force[me + 3] = pot * gm;
  }
}

/* The same field in FP64 from the same FP32 mesh: the check that pins the
 * FP32 restatement's arithmetic (a dropped term or a transposed tensor entry
 * shows as an O(1) difference, rounding as ~1e-6). */
void cco_cell_f64(const float* fine, const float* coarse, const float* farA, const float* farT0, const float* farT1,
                  double eps2, double gm, double* force) {
  for (int my = 0; my < 512; my++) for (int mx = 0; mx < 512; mx++) {
    size_t me = ((size_t)my * 512 + mx) * 4;
    int c[3] = { mx & 63, my & 63, (my >> 6) * 8 + (mx >> 6) };
    const float* mc = fine + me;
    for (int a = 0; a < 4; a++) force[me + a] = 0.0;
    if (mc[3] <= 0.0f) continue;
    double x[3] = { (double)mc[0] / mc[3], (double)mc[1] / mc[3], (double)mc[2] / mc[3] };
    int cc[3] = { c[0] >> 2, c[1] >> 2, c[2] >> 2 };
    double acc[3] = { 0, 0, 0 }, pot = 0.0;
    for (int fz = 0; fz < 64; fz++) for (int fy = 0; fy < 64; fy++) for (int fx = 0; fx < 64; fx++) {
      if (abs((fx >> 2) - cc[0]) > 2 || abs((fy >> 2) - cc[1]) > 2 || abs((fz >> 2) - cc[2]) > 2) continue;
      if (fx == c[0] && fy == c[1] && fz == c[2]) continue;
      const float* mf = fine + (size_t)fine_at(fx, fy, fz) * 4;
      if (mf[3] <= 0.0f) continue;
      double r[3] = { (double)mf[0] / mf[3] - x[0], (double)mf[1] / mf[3] - x[1], (double)mf[2] / mf[3] - x[2] };
      double inv = 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + eps2);
      for (int a = 0; a < 3; a++) acc[a] += r[a] * mf[3] * inv * inv * inv;
      pot -= mf[3] * inv;
    }
    size_t ca = (size_t)coarse_at(cc[0], cc[1], cc[2]) * 4;
    const float *mb = coarse + ca, *fa = farA + ca, *t0 = farT0 + ca, *t1 = farT1 + ca;
    double d[3] = { x[0] - (double)mb[0] / mb[3], x[1] - (double)mb[1] / mb[3], x[2] - (double)mb[2] / mb[3] };
    /* symmetric tidal tensor T = [[Txx Txy Txz],[Txy Tyy Tyz],[Txz Tyz Tzz]] */
    double T[3][3] = { { t0[0], t0[3], t1[0] }, { t0[3], t0[1], t1[1] }, { t1[0], t1[1], t0[2] } };
    double Td[3];
    for (int a = 0; a < 3; a++) Td[a] = T[a][0] * d[0] + T[a][1] * d[1] + T[a][2] * d[2];
    for (int a = 0; a < 3; a++) acc[a] += fa[a] + Td[a];
    pot += fa[3] - (fa[0] * d[0] + fa[1] * d[1] + fa[2] * d[2]) - 0.5 * (d[0] * Td[0] + d[1] * Td[1] + d[2] * Td[2]);
    force[me] = acc[0] * gm; force[me + 1] = acc[1] * gm; force[me + 2] = acc[2] * gm; force[me + 3] = pot * gm;
  }
}

void cco_pp_gather(const float* pos, const float* mat, const float* index, const cco_box* b, float* list, float* box) {
  for (int s = 0; s < 8192; s++) {
    float fi = index[s];
    if (fi < 0.0f) { list[s * 4] = list[s * 4 + 1] = list[s * 4 + 2] = list[s * 4 + 3] = 0.0f; box[s] = -1.0f; continue; }
    int i = (int)fi;
    const float* p = pos + (size_t)i * 4;
    float g[3];
    box[s] = (float)cco_box_of(b, p, g);
    list[s * 4] = p[0]; list[s * 4 + 1] = p[1]; list[s * 4 + 2] = p[2]; list[s * 4 + 3] = mat[i];
  }
}

static inline float smoothstepf(float e0, float e1, float x) {
  float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

void cco_pp_force(const float* list, const float* box, int count, float gm, float cut, float a2, const float* tab,
                  float* force) {
  int share = (count + 3) >> 2;
  float cut2 = cut * cut, a = sqrtf(a2);
  for (int fy = 0; fy < 256; fy++) for (int fx = 0; fx < 128; fx++) {
    float* o = force + ((size_t)fy * 128 + fx) * 4;
    int slice = fy >> 6, my = fy & 63, myIdx = my * 128 + fx;
    int j0 = slice * share, j1 = imin(j0 + share, count);
    const float* P = list + (size_t)myIdx * 4;
    o[0] = o[1] = o[2] = o[3] = 0.0f;
    if (myIdx >= count || P[3] == 0.0f) continue;
    float myBox = box[myIdx];
    float acc[3] = { 0, 0, 0 }, pot = 0.0f;
    for (int j = j0; j < j1; j++) {
      if (j == myIdx) continue;
      const float* Q = list + (size_t)j * 4;
      float d[3] = { Q[0] - P[0], Q[1] - P[1], Q[2] - P[2] };
      float r2 = dot3(d, d);
      if (r2 >= cut2) continue;
      float r = sqrtf(r2);
      float x = clampf(r / cut * 64.0f - 0.5f, 0.0f, 63.0f);
      int k = (int)x, k1 = imin(k + 1, 63);
      float t = x - (float)k;
      float invN = rsq(r2 + a2);
      float taper = smoothstepf(cut, 0.8f * cut, r);
      float same = (myBox >= 0.0f && box[j] == myBox) ? 1.0f : 0.0f;
      float Tx = mixf(tab[k * 2], tab[k1 * 2], t) * same, Ty = mixf(tab[k * 2 + 1], tab[k1 * 2 + 1], t) * same;
      float fn = invN * invN * invN * r - Tx / fmaxf_(r2, a2), mj = Q[3];
      float s = mj * fn * taper, rr = fmaxf_(r, 1e-6f);
      for (int c = 0; c < 3; c++) acc[c] += d[c] / rr * s;
      pot += mj * (-invN - Ty / fmaxf_(r, a)) * taper;
    }
    o[0] = acc[0] * gm; o[1] = acc[1] * gm; o[2] = acc[2] * gm; o[3] = pot * gm;
  }
}

static inline uint32_t seat_of(int k, const uint32_t* sa, const uint32_t* sb) {
  return (k & 1) == 0 ? sa[k >> 1] : sb[k >> 1];
}

static void mass_at(const float p[3], const float* mono, float gm, float eps2, float out[3]) {
  float r[3];
  for (int a = 0; a < 3; a++) r[a] = mono[a] / fmaxf_(mono[3], 1e-6f) - p[a];
  float inv = rsq(dot3(r, r) + eps2);
  float s = gm * mono[3] * inv * inv * inv;
  for (int a = 0; a < 3; a++) out[a] = r[a] * s;
}
static float pot_at(const float p[3], const float* mono, float gm, float eps2) {
  float r[3];
  for (int a = 0; a < 3; a++) r[a] = mono[a] / fmaxf_(mono[3], 1e-6f) - p[a];
  return -gm * mono[3] * rsq(dot3(r, r) + eps2);
}
static void cross_at(const float* mo, const float* mm, float gm, float eps2, float out[3]) {
  float d[3];
  for (int a = 0; a < 3; a++) d[a] = mo[a] / fmaxf_(mo[3], 1e-6f) - mm[a] / fmaxf_(mm[3], 1e-6f);
  float inv = rsq(dot3(d, d) + eps2);
  float s = gm * mo[3] * inv * inv * inv;
  for (int a = 0; a < 3; a++) out[a] = d[a] * s;
}
static inline float mod4(float x) { return x - 4.0f * floorf(x / 4.0f); }

void cco_sim(const cco_sim_u* u, const float* pos, const float* vel, const float* mat, const float* q,
             const uint32_t* seatA, const uint32_t* seatB, const float* force0, const float* mono0,
             const float* force1, const float* mono1, const float* ppSlot, const float* ppForce,
             float* oPos, float* oVel, float* oAux, float* oQ) {
  int N = u->W * u->H;
  const cco_grid* gr = &u->grid;
  for (int me = 0; me < N; me++) {
    const float* P = pos + (size_t)me * 4;
    const float* V = vel + (size_t)me * 4;
    float p[3] = { P[0], P[1], P[2] }, v[3] = { V[0], V[1], V[2] };
    float body = P[3];
    float ri = mat[me];
    float ci = ri * (mod4(body) < 0.5f ? u->cpIron : 1.0f);
    float Ti = V[3];
    float push[3] = { 0, 0, 0 }, nb = 0.0f, lost = 0.0f, flow = 0.0f;
    int c[3];
    for (int a = 0; a < 3; a++) c[a] = (int)floorf(p[a] * u->invCell);
    float c0[3] = { (float)c[0] * u->cell, (float)c[1] * u->cell, (float)c[2] * u->cell };
    for (int dz = -1; dz <= 1; dz++) for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
      float dd[3] = { (float)dx, (float)dy, (float)dz }, lo[3], qq[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = c0[a] + dd[a] * u->cell;
        qq[a] = clampf(p[a], lo[a], lo[a] + u->cell) - p[a];
      }
      if (dot3(qq, qq) > u->link2) continue;
      int ax, ay;
      cco_hash_at(gr, c[0] + dx, c[1] + dy, c[2] + dz, &ax, &ay);
      size_t t = ((size_t)ay * gr->AW + ax) * 4;
      const uint32_t *sa = seatA + t, *sb = seatB + t;
      for (int k = 0; k < 8; k++) {
        uint32_t sv = seat_of(k, sa, sb);
        if (sv & 0x80000000u) break;
        int j = (int)(sv & ((1u << u->grid.IB) - 1u));
        if (j == me) continue;
        int B = u->grid.IB;
        float sp[3] = { ((float)((sv >> B) & 15u) + 0.5f) / 15.0f, ((float)((sv >> (B + 4)) & 15u) + 0.5f) / 15.0f,
                        ((float)((sv >> (B + 8)) & 15u) + 0.5f) / 15.0f };
        float dq[3];
        for (int a = 0; a < 3; a++) dq[a] = lo[a] + sp[a] * u->cell - p[a];
        if (dot3(dq, dq) > u->reach2) continue;
        const float* Pj = pos + (size_t)j * 4;
        float d[3] = { Pj[0] - p[0], Pj[1] - p[1], Pj[2] - p[2] };
        float r2 = dot3(d, d);
        if (r2 < u->link2) {
          float r = sqrtf(r2);
          float rr = fmaxf_(r, 1e-6f);
          float n[3] = { d[0] / rr, d[1] / rr, d[2] / rr };
          float ov = u->touch - r;
          if (ov > 0.0f) {
            const float* Vj = vel + (size_t)j * 4;
            float rj = mat[j];
            float cij = u->c * sqrtf(2.0f * ri * rj / (ri + rj));
            float dv[3] = { Vj[0] - v[0], Vj[1] - v[1], Vj[2] - v[2] };
            float vn = dot3(dv, n);
            float fs = u->k * ov - cij * vn, fn = fmaxf_(fs, 0.0f), on = fs < 0.0f ? 0.0f : 1.0f;
            for (int a = 0; a < 3; a++) push[a] -= n[a] * fn;
            lost += cij * vn * vn * on + u->k * ov * vn * (1.0f - on);
            float cj = rj * (mod4(Pj[3]) < 0.5f ? u->cpIron : 1.0f);
            flow += (Vj[3] - Ti) * (2.0f * ci * cj / (ci + cj));
          }
          nb += 1.0f;
        }
      }
    }
    float grav[3] = { 0, 0, 0 };
    if (u->field) {
      for (int a = 0; a < 3; a++) grav[a] = u->gField[a];
    } else {
      float g[3];
      int box = cco_box_of(&u->box, p, g);
      const float* F = box == 1 ? force1 : force0;
      if (box >= 0) {
        float f[3], t[3];
        int ci0[3];
        for (int a = 0; a < 3; a++) { f[a] = floorf(g[a]); t[a] = g[a] - f[a]; ci0[a] = (int)f[a]; }
        for (int k = 0; k < 8; k++) {
          int o[3] = { k & 1, (k >> 1) & 1, k >> 2 };
          float w3[3];
          for (int a = 0; a < 3; a++) w3[a] = mixf(1.0f - t[a], t[a], (float)o[a]);
          const float* fv = F + (size_t)fine_at(imin(ci0[0] + o[0], 63), imin(ci0[1] + o[1], 63), imin(ci0[2] + o[2], 63)) * 4;
          float w = w3[0] * w3[1] * w3[2];
          for (int a = 0; a < 3; a++) grav[a] += fv[a] * w;
        }
        const float* sc = F + (size_t)512 * 512 * 4;
        for (int a = 0; a < 3; a++) grav[a] -= sc[a] * u->gm;
        if (u->box.boxB > 0.5f) {
          float cr[3];
          if (box == 0) cross_at(mono1, mono0, u->gm, u->eps2, cr); else cross_at(mono0, mono1, u->gm, u->eps2, cr);
          for (int a = 0; a < 3; a++) grav[a] += cr[a];
        }
      } else {
        mass_at(p, mono0, u->gm, u->eps2, grav);
        if (u->box.boxB > 0.5f) {
          float mb[3];
          mass_at(p, mono1, u->gm, u->eps2, mb);
          for (int a = 0; a < 3; a++) grav[a] += mb[a];
        }
      }
      float seat = ppSlot[me];
      if (seat >= 0.0f) {
        int si = (int)seat, sx = si & 127, sy = si >> 7;
        for (int s = 0; s < 4; s++) {
          const float* pf = ppForce + ((size_t)(sy + 64 * s) * 128 + sx) * 4;
          for (int a = 0; a < 3; a++) grav[a] += pf[a];
        }
      }
    }
    /* SIM_POST (src/glsl-sim.js:494-516) */
    float sc = u->invM / ri;
    for (int a = 0; a < 3; a++) v[a] += push[a] * sc * u->dt + grav[a] * u->gdt;
    int tgt = body < 3.5f;
    const float* sp = tgt ? u->spin0 : u->spin1;
    const float* cn = tgt ? u->cen0 : u->cen1;
    float arm[3] = { p[0] - cn[0], p[1] - cn[1], p[2] - cn[2] };
    float vr[3] = { sp[1] * arm[2] - sp[2] * arm[1], sp[2] * arm[0] - sp[0] * arm[2], sp[0] * arm[1] - sp[1] * arm[0] };
    if (u->settle > 0.5f) {
      float dv[3];
      for (int a = 0; a < 3; a++) dv[a] = (v[a] - vr[a]) * u->damp;
      float spd = sqrtf(dot3(dv, dv));
      float f = fminf_(1.0f, u->creep / fmaxf_(spd, 1e-9f));
      for (int a = 0; a < 3; a++) v[a] = vr[a] + dv[a] * f;
    }
    if (u->kick > 0.5f) {
      const float* bk = tgt ? u->bulk0 : u->bulk1;
      for (int a = 0; a < 3; a++) v[a] = bk[a] + vr[a];
    }
    float Q = u->settle > 0.5f ? 0.0f : q[me] + (lost + flow * u->kappa / u->tK) * u->dt;
    if (u->frozenOn && fabsf(body - u->frozen) < 0.5f) v[0] = v[1] = v[2] = 0.0f;
    for (int a = 0; a < 3; a++) p[a] += v[a] * u->dt;
    float* op = oPos + (size_t)me * 4;
    float* ov = oVel + (size_t)me * 4;
    float* oa = oAux + (size_t)me * 4;
    op[0] = p[0]; op[1] = p[1]; op[2] = p[2]; op[3] = body;
    ov[0] = v[0]; ov[1] = v[1]; ov[2] = v[2]; ov[3] = u->tK * Q / ci;
    oa[0] = grav[0]; oa[1] = grav[1]; oa[2] = grav[2]; oa[3] = nb;
    oQ[me] = Q;
  }
}

static void mat3_mul(const float* m, const float* v, float* out) {   /* column-major */
  for (int a = 0; a < 3; a++) out[a] = m[a] * v[0] + m[3 + a] * v[1] + m[6 + a] * v[2];
}

void cco_rigid(int N, const float* pos, const float* vel, const float* aux, const float* q, const float bulk0[3],
               const float bulk1[3], const float cen0[3], const float cen1[3], const float spin0[3], const float spin1[3],
               const float rot0[9], const float rot1[9], float dt, float* oPos, float* oVel, float* oAux, float* oQ) {
  for (int i = 0; i < N; i++) {
    const float* P = pos + (size_t)i * 4;
    int tgt = P[3] < 3.5f;
    const float *bulk = tgt ? bulk0 : bulk1, *cen = tgt ? cen0 : cen1, *sp = tgt ? spin0 : spin1, *rot = tgt ? rot0 : rot1;
    float rel[3] = { P[0] - cen[0], P[1] - cen[1], P[2] - cen[2] }, arm[3];
    mat3_mul(rot, rel, arm);
    for (int a = 0; a < 3; a++) oPos[i * 4 + a] = cen[a] + arm[a] + bulk[a] * dt;
    oPos[i * 4 + 3] = P[3];
    float cr[3] = { sp[1] * arm[2] - sp[2] * arm[1], sp[2] * arm[0] - sp[0] * arm[2], sp[0] * arm[1] - sp[1] * arm[0] };
    for (int a = 0; a < 3; a++) oVel[i * 4 + a] = bulk[a] + cr[a];
    oVel[i * 4 + 3] = vel[i * 4 + 3];
    for (int a = 0; a < 4; a++) oAux[i * 4 + a] = aux[i * 4 + a];
    oQ[i] = q[i];
  }
}

void cco_perm(int N, const float* perm, const float* pos, const float* vel, const float* aux, const float* q,
              float* oPos, float* oVel, float* oAux, float* oQ) {
  for (int i = 0; i < N; i++) {
    int s = (int)perm[i];
    memcpy(oPos + (size_t)i * 4, pos + (size_t)s * 4, 16);
    memcpy(oVel + (size_t)i * 4, vel + (size_t)s * 4, 16);
    memcpy(oAux + (size_t)i * 4, aux + (size_t)s * 4, 16);
    oQ[i] = q[s];
  }
}

void cco_couple(int N, const float* args, const float* pos, const float* vel, const float* aux, const float* mat,
                float* oPos, float* oVel, float* oAux, float* oQ) {
  const float *centre = args;
  float vimp = args[3], aimp = args[4], c1 = args[5], invMu = args[6], pexp = args[7], stall = args[8], z = args[9];
  float heat = args[10], tK = args[11], frozen = args[12], ambient = args[13], melt = args[14];
  for (int i = 0; i < N; i++) {
    const float *P = pos + (size_t)i * 4, *V = vel + (size_t)i * 4;
    float w = P[3], ri = mat[i];
    memcpy(oPos + (size_t)i * 4, P, 16);
    memcpy(oAux + (size_t)i * 4, aux + (size_t)i * 4, 16);
    if (fabsf(w - frozen) < 0.5f) {
      oVel[i * 4] = oVel[i * 4 + 1] = oVel[i * 4 + 2] = 0.0f; oVel[i * 4 + 3] = ambient; oQ[i] = 0.0f;
      continue;
    }
    float R[3] = { P[0] - centre[0], P[1] - centre[1], P[2] - centre[2] };
    float r = fmaxf_(sqrtf(dot3(R, R)), 1e-4f);
    float v = c1 * vimp * powf(r / aimp, -invMu);
    v = fminf_(v, vimp);
    v *= powf(fmaxf_(0.0f, 1.0f - r / stall), pexp);
    float n[3] = { R[0] / r, R[1] / r, R[2] / r }, axis[3] = { 0.0f, -1.0f, 0.0f };
    float ct = clampf(dot3(n, axis), -1.0f, 1.0f);
    float st = sqrtf(fmaxf_(1.0f - ct * ct, 1e-8f));
    float eth[3];
    for (int a = 0; a < 3; a++) eth[a] = (ct * n[a] - axis[a]) / st;
    float th = acosf(ct);
    float f = (z - 2.0f) * tanf(0.5f * th);
    float d[3];
    for (int a = 0; a < 3; a++) d[a] = n[a] + f * eth[a];
    float inv = rsq(dot3(d, d));
    float vv[3];
    for (int a = 0; a < 3; a++) vv[a] = d[a] * inv * v;
    int imp = w > 3.5f;
    if (imp) for (int a = 0; a < 3; a++) vv[a] = mixf(vv[a], V[a], 0.5f);
    float speed = sqrtf(dot3(vv, vv));
    float T = ambient + heat * 0.5f * speed * speed;
    oVel[i * 4] = vv[0]; oVel[i * 4 + 1] = vv[1]; oVel[i * 4 + 2] = vv[2]; oVel[i * 4 + 3] = T;
    oQ[i] = T * ri / fmaxf_(tK, 1e-9f);
    oPos[i * 4 + 3] = imp ? melt : (T > 1500.0f ? melt : w);
  }
}

void cco_diag(const cco_sim_u* u, const float* pos, const float* q, const uint32_t* seatA, const uint32_t* seatB,
              const float* force0, const float* mono0, const float* force1, const float* mono1, const float* ppSlot,
              const float* ppForce, float* out) {
  int N = u->W * u->H;
  const cco_grid* gr = &u->grid;
  for (int me = 0; me < N; me++) {
    const float* P = pos + (size_t)me * 4;
    float p[3] = { P[0], P[1], P[2] };
    int ax, ay;
    cco_hash_at(gr, (int)floorf(p[0] * u->invCell), (int)floorf(p[1] * u->invCell), (int)floorf(p[2] * u->invCell), &ax, &ay);
    size_t t = ((size_t)ay * gr->AW + ax) * 4;
    float seated = 0.0f, taken = 0.0f;
    for (int k = 0; k < 8; k++) {
      uint32_t sv = seat_of(k, seatA + t, seatB + t);
      if (sv & 0x80000000u) break;
      taken += 1.0f;
      if ((int)(sv & ((1u << u->grid.IB) - 1u)) == me) seated = 1.0f;
    }
    float phi = 0.0f, off = 0.0f, g[3];
    int box = cco_box_of(&u->box, p, g);
    if (box >= 0) {
      const float* F = box == 1 ? force1 : force0;
      float f[3], tt[3];
      int ci0[3];
      for (int a = 0; a < 3; a++) { f[a] = floorf(g[a]); tt[a] = g[a] - f[a]; ci0[a] = (int)f[a]; }
      for (int k = 0; k < 8; k++) {
        int o[3] = { k & 1, (k >> 1) & 1, k >> 2 };
        float w3[3];
        for (int a = 0; a < 3; a++) w3[a] = mixf(1.0f - tt[a], tt[a], (float)o[a]);
        phi += F[(size_t)fine_at(imin(ci0[0] + o[0], 63), imin(ci0[1] + o[1], 63), imin(ci0[2] + o[2], 63)) * 4 + 3] * (w3[0] * w3[1] * w3[2]);
      }
      if (u->box.boxB > 0.5f) phi += box == 0 ? pot_at(p, mono1, u->gm, u->eps2) : pot_at(p, mono0, u->gm, u->eps2);
    } else {
      phi = pot_at(p, mono0, u->gm, u->eps2);
      if (u->box.boxB > 0.5f) phi += pot_at(p, mono1, u->gm, u->eps2);
      off = 2.0f;
    }
    float seat = ppSlot[me];
    if (seat >= 0.0f) {
      int si = (int)seat, sx = si & 127, sy = si >> 7;
      for (int s = 0; s < 4; s++) phi += ppForce[((size_t)(sy + 64 * s) * 128 + sx) * 4 + 3];
    }
    out[me * 4] = phi; out[me * 4 + 1] = seated + off; out[me * 4 + 2] = taken; out[me * 4 + 3] = q[me];
  }
}

void cco_morton(const float* pos, int N, const double core[3], int32_t* keys, float* perm, int32_t* into) {
  const int KB = 1 << 21;
  int32_t* bins = (int32_t*)calloc((size_t)KB + 1, sizeof(int32_t));
  for (int i = 0; i < N; i++) {
    int m[3];
    for (int a = 0; a < 3; a++) {
      double v = floor((((double)pos[i * 4 + a] - core[a]) + 6.0) / 12.0 * 128.0);
      v = v < 0.0 ? 0.0 : v;
      v = v > 127.0 ? 127.0 : v;
      m[a] = (int)v;
    }
    int mk = 0;
    for (int k = 0; k < 7; k++)
      mk |= ((m[0] >> k) & 1) << (3 * k) | ((m[1] >> k) & 1) << (3 * k + 1) | ((m[2] >> k) & 1) << (3 * k + 2);
    keys[i] = mk;
    bins[mk + 1]++;
  }
  for (int k = 0; k < KB; k++) bins[k + 1] += bins[k];
  for (int i = 0; i < N; i++) { int dst = bins[keys[i]]++; perm[dst] = (float)i; into[i] = dst; }
  free(bins);
}

void cco_books(int N, double mref, const float* pos, const float* vel, const float* mat, const float* diag, double* out) {
  double Mt = 0, px = 0, py = 0, pz = 0, Jx = 0, Jy = 0, Jz = 0, ke = 0, pe = 0, Q = 0, ironAll = 0;
  int unseated = 0, full = 0;
  for (int i = 0; i < N; i++) {
    double m = mref * mat[i];
    double bx = pos[i * 4], by = pos[i * 4 + 1], bz = pos[i * 4 + 2];
    double bu = vel[i * 4], bv = vel[i * 4 + 1], bw = vel[i * 4 + 2];
    double w = pos[i * 4 + 3];
    Mt += m; px += m * bu; py += m * bv; pz += m * bw;
    if (w - 4 * floor(w / 4) < 0.5) ironAll += m;
    Jx += m * (by * bw - bz * bv); Jy += m * (bz * bu - bx * bw); Jz += m * (bx * bv - by * bu);
    ke += 0.5 * m * (bu * bu + bv * bv + bw * bw);
    double flag = diag[i * 4 + 1];
    int off = flag >= 2;
    pe += (off ? 1 : 0.5) * m * diag[i * 4];
    Q += 0.5 * diag[i * 4 + 3];
    if (flag - (off ? 2 : 0) < 0.5) unseated++;
    if (diag[i * 4 + 2] >= 7.5) full++;
  }
  out[0] = Mt; out[1] = px; out[2] = py; out[3] = pz; out[4] = Jx; out[5] = Jy; out[6] = Jz;
  out[7] = ke; out[8] = pe; out[9] = Q; out[10] = (double)unseated / N; out[11] = (double)full / N;
  out[12] = ironAll / Mt; out[13] = sqrt(px * px + py * py + pz * pz) / Mt;
}

uint32_t cco_mulberry_next(uint32_t* state, double* out) {
  uint32_t s = *state + 0x6D2B79F5u;
  *state = s;
  uint32_t t = (s ^ (s >> 15)) * (1u | s);
  t = (t + (t ^ (t >> 7)) * (61u | t)) ^ t;
  uint32_t r = t ^ (t >> 14);
  if (out) *out = (double)r / 4294967296.0;
  return r;
}

/* meshPairTable: the pair force and potential the mesh gives two unit
 * particles at separation r (deposit, near field at node centres of mass with
 * softening 0.5 FH^2, own node skipped, CIC gather), less the first
 * particle's self-force; 80 random placements per bin, Mulberry32 seed 777,
 * nodes kept in first-touch order as the reference's Map is. */
typedef struct { long key; double m, x, y, z, fx, fy, fz, pot; } pt_node;
#define PT_FL 6.0
#define PT_FH (2.0 * PT_FL / 64.0)
static void pt_corners(const double* p, long keys[8], double w[8]) {
  double g[3], fl[3], t[3];
  for (int a = 0; a < 3; a++) { g[a] = (p[a] + PT_FL) / PT_FH - 0.5; fl[a] = floor(g[a]); t[a] = g[a] - fl[a]; }
  for (int k = 0; k < 8; k++) {
    int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
    long ix = (long)fl[0] + ox, iy = (long)fl[1] + oy, iz = (long)fl[2] + oz;
    keys[k] = ((ix + 4096) * 8192 + (iy + 4096)) * 8192 + (iz + 4096);
    w[k] = (ox ? t[0] : 1 - t[0]) * (oy ? t[1] : 1 - t[1]) * (oz ? t[2] : 1 - t[2]);
  }
}
static void pt_field(const double (*parts)[3], int np, const double* at, double out[4]) {
  pt_node nodes[16];
  int nn = 0;
  const double eps2 = 0.5 * PT_FH * PT_FH;
  for (int i = 0; i < np; i++) {
    long keys[8]; double w[8];
    pt_corners(parts[i], keys, w);
    for (int k = 0; k < 8; k++) {
      int f = -1;
      for (int q = 0; q < nn; q++) if (nodes[q].key == keys[k]) { f = q; break; }
      if (f < 0) { f = nn++; memset(&nodes[f], 0, sizeof(pt_node)); nodes[f].key = keys[k]; }
      nodes[f].m += w[k]; nodes[f].x += w[k] * parts[i][0]; nodes[f].y += w[k] * parts[i][1]; nodes[f].z += w[k] * parts[i][2];
    }
  }
  for (int q = 0; q < nn; q++) { nodes[q].x /= nodes[q].m; nodes[q].y /= nodes[q].m; nodes[q].z /= nodes[q].m; }
  for (int c = 0; c < nn; c++) for (int o = 0; o < nn; o++) {
    if (o == c) continue;
    double rx = nodes[o].x - nodes[c].x, ry = nodes[o].y - nodes[c].y, rz = nodes[o].z - nodes[c].z;
    double inv = 1 / sqrt(rx * rx + ry * ry + rz * rz + eps2), inv3 = inv * inv * inv;
    nodes[c].fx += rx * nodes[o].m * inv3; nodes[c].fy += ry * nodes[o].m * inv3; nodes[c].fz += rz * nodes[o].m * inv3;
    nodes[c].pot -= nodes[o].m * inv;
  }
  out[0] = out[1] = out[2] = out[3] = 0;
  long keys[8]; double w[8];
  pt_corners(at, keys, w);
  for (int k = 0; k < 8; k++) {
    int f = -1;
    for (int q = 0; q < nn; q++) if (nodes[q].key == keys[k]) { f = q; break; }
    if (f < 0) continue;
    out[0] += w[k] * nodes[f].fx; out[1] += w[k] * nodes[f].fy; out[2] += w[k] * nodes[f].fz; out[3] += w[k] * nodes[f].pot;
  }
}
void cco_mesh_pair_table(int bins, double cut, float* f, float* p) {
  uint32_t st = 777;
  for (int b = 0; b < bins; b++) {
    double r = (b + 0.5) / bins * cut, sf = 0, sp = 0;
    for (int s = 0; s < 80; s++) {
      double r0, r1, r2, r3, r4;
      cco_mulberry_next(&st, &r0); cco_mulberry_next(&st, &r1); cco_mulberry_next(&st, &r2);
      double pp[3] = { (r0 - 0.5) * PT_FH, (r1 - 0.5) * PT_FH, (r2 - 0.5) * PT_FH };
      cco_mulberry_next(&st, &r3);
      double z = 2 * r3 - 1;
      cco_mulberry_next(&st, &r4);
      double ph = r4 * 6.2831853, rr = sqrt(1 - z * z);
      double u[3] = { rr * cos(ph), rr * sin(ph), z };
      double qq[3] = { pp[0] + u[0] * r, pp[1] + u[1] * r, pp[2] + u[2] * r };
      double both[4], alone[4];
      double two[2][3] = { { pp[0], pp[1], pp[2] }, { qq[0], qq[1], qq[2] } };
      pt_field((const double (*)[3])two, 2, pp, both);
      pt_field((const double (*)[3])two, 1, pp, alone);
      sf += (both[0] - alone[0]) * u[0] + (both[1] - alone[1]) * u[1] + (both[2] - alone[2]) * u[2];
      sp += both[3] - alone[3];
    }
    f[b] = (float)(sf / 80 * r * r);
    p[b] = (float)(sp / 80 * r);
  }
}

/* The FP64 direct sum the mesh and pairwise gravity approximate: Newton
 * softened at the particle radius (the pairwise pass's own softening,
 * src/glsl-sim.js:735-751), every pair, acceleration and potential times G m.
 * O(N^2): for the 16,384-particle states only. */
void cco_direct_f64(int N, const float* pos, const float* mat, double gm, double a2, double* out) {
  for (int i = 0; i < N; i++) {
    double px = pos[i * 4], py = pos[i * 4 + 1], pz = pos[i * 4 + 2];
    double ax = 0, ay = 0, az = 0, pot = 0;
    for (int j = 0; j < N; j++) {
      if (j == i) continue;
      double dx = pos[j * 4] - px, dy = pos[j * 4 + 1] - py, dz = pos[j * 4 + 2] - pz;
      double r2 = dx * dx + dy * dy + dz * dz + a2, inv = 1.0 / sqrt(r2), s = mat[j] * inv * inv * inv;
      ax += dx * s; ay += dy * s; az += dz * s; pot -= mat[j] * inv;
    }
    out[i * 4] = ax * gm; out[i * 4 + 1] = ay * gm; out[i * 4 + 2] = az * gm; out[i * 4 + 3] = pot * gm;
  }
}

/* ---- friends-of-friends linking (src/nbody.js:647-684) ----
 * The reference's own sequence: cells of width 2.4a hashed into 2^20 buckets
 * (products in FP64, each operand of ^ taken to int32 as JavaScript does),
 * a counting sort in index order, then for i ascending the 27 buckets dz, dy,
 * dx from -1 to 1 and each bucket's members in order, uniting every pair
 * j > i closer than 2.4a with the reference's union-find (uf[ri] = rj).
 * Outputs: root[i] = that union-find's root, label[i] = the smallest index in
 * i's group, e2[] = the touching pairs' squared distances in pair order (the
 * count is returned; e2 may be NULL or short, then only counted), the elastic
 * energy summed as the reference sums it, and when linkOff is given every
 * linked partner in order (linkOff[N+1], partner[] of room `pcap`). */
static int32_t js_int32(double d) {
  if (!(d == d) || isinf(d)) return 0;
  if (d > -9.2e18 && d < 9.2e18) return (int32_t)(uint32_t)(uint64_t)(int64_t)d;
  double m = fmod(trunc(d), 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (int32_t)(uint32_t)m;
}
static int fof_hash(double x, double y, double z) {
  return (js_int32(x * 73856093.0) ^ js_int32(y * 19349663.0) ^ js_int32(z * 83492791.0)) & ((1 << 20) - 1);
}
static int uf_find(int* uf, int x) {
  while (uf[x] != x) { uf[x] = uf[uf[x]]; x = uf[x]; }
  return x;
}
long cco_fof_link(int N, const float* P, double a, double touch, double spring, int32_t* root, int32_t* label,
                  double* e2out, long e2cap, double* elastic, int32_t* linkOff, int32_t* partner, long pcap) {
  const int T = 1 << 20;
  double L = 2.4 * a, L2 = L * L, inv = 1 / L, touch2 = touch * touch, el = 0;
  int32_t *cx = malloc(sizeof(int32_t) * N), *cy = malloc(sizeof(int32_t) * N), *cz = malloc(sizeof(int32_t) * N);
  int32_t *key = malloc(sizeof(int32_t) * N), *cnt = calloc((size_t)T + 1, sizeof(int32_t));
  int32_t *fill = malloc(sizeof(int32_t) * T), *order = malloc(sizeof(int32_t) * N), *uf = malloc(sizeof(int32_t) * N);
  for (int i = 0; i < N; i++) {
    double qx = floor((double)P[i * 4] * inv), qy = floor((double)P[i * 4 + 1] * inv), qz = floor((double)P[i * 4 + 2] * inv);
    cx[i] = js_int32(qx); cy[i] = js_int32(qy); cz[i] = js_int32(qz);
    key[i] = fof_hash(qx, qy, qz);
    cnt[key[i] + 1]++;
  }
  for (int k = 0; k < T; k++) cnt[k + 1] += cnt[k];
  memcpy(fill, cnt, sizeof(int32_t) * T);
  for (int i = 0; i < N; i++) order[fill[key[i]]++] = i;
  for (int i = 0; i < N; i++) uf[i] = i;
  long nt = 0, np = 0;
  for (int i = 0; i < N; i++) {
    double px = P[i * 4], py = P[i * 4 + 1], pz = P[i * 4 + 2];
    if (linkOff) linkOff[i] = (int32_t)np;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int k = fof_hash((double)cx[i] + dx, (double)cy[i] + dy, (double)cz[i] + dz);
          for (int s = cnt[k]; s < cnt[k + 1]; s++) {
            int j = order[s];
            if (j <= i) continue;
            double ex = (double)P[j * 4] - px, ey = (double)P[j * 4 + 1] - py, ez = (double)P[j * 4 + 2] - pz;
            double e2 = ex * ex + ey * ey + ez * ez;
            if (e2 < L2) {
              int ri = uf_find(uf, i), rj = uf_find(uf, j);
              if (ri != rj) uf[ri] = rj;
              if (linkOff && np < pcap) partner[np] = j;
              np++;
              if (e2 < touch2) {
                double ov = touch - sqrt(e2);
                el += 0.5 * spring * ov * ov;
                if (e2out && nt < e2cap) e2out[nt] = e2;
                nt++;
              }
            }
          }
        }
  }
  if (linkOff) linkOff[N] = (int32_t)np;
  for (int i = 0; i < N; i++) root[i] = uf_find(uf, i);
  /* smallest member of each root's group */
  int32_t* small = malloc(sizeof(int32_t) * N);
  for (int i = 0; i < N; i++) small[i] = -1;
  for (int i = 0; i < N; i++) if (small[root[i]] < 0) small[root[i]] = i;
  for (int i = 0; i < N; i++) label[i] = small[root[i]];
  *elastic = el;
  free(small); free(cx); free(cy); free(cz); free(key); free(cnt); free(fill); free(order); free(uf);
  return nt;
}
