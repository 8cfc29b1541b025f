/* CPU oracle for the cosmic-collisions simulation step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark may call it; the product
 * (the HIP library) never links or calls it.
 *
 * Each function is a plain FP32 (or, where the reference computes on the
 * host, FP64) restatement of one reference pass, in the reference's own
 * texture layouts so that outputs compare texel for texel:
 *   particle arrays  : index i = y*W + x of the W x H state textures
 *   pos/vel/aux      : float[N*4];  q, mat: float[N]
 *   seat atlases     : uint32[AW*AH*4] (the RGBA32F bit patterns), two of them
 *   fine/force mesh  : float[513*512*4] (8x8 atlas of 64x64 slices + scratch row)
 *   coarse/far       : float[64*64*4];  mono: float[4]
 *   loose list       : float[128*64*4], box float[128*64], force float[128*256*4]
 * Citations are to the reference (Gaploid/cosmic-collisions) source files.
 */
#ifndef CC_ORACLE_H
#define CC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int G, GShift, GMask, SX, SXShift, SXMask, AW, AH;
  int IB;                 /* seat word index bits: 18 (the reference's); 19 above 262,144 particles */
} cco_grid;

typedef struct {          /* the two mesh boxes, as the shaders see them */
  float invH;             /* 1 / FH */
  float LA[3], LB[3];     /* half-width less centre */
  float boxB;             /* 1: the second box is up */
} cco_box;

typedef struct {
  int W, H;
  float invCell, cell, dt, gdt, touch, link2, reach2, k, c, invM, damp, creep, tK, kappa, cpIron;
  float settle, kick;
  float bulk0[3], bulk1[3], cen0[3], cen1[3], spin0[3], spin1[3];
  cco_grid grid;
  cco_box box;
  float gm, eps2;         /* mesh gravity: G*m and the monopole softening (FH^2) */
  int field;              /* 1: constant gravity gField instead of the mesh */
  float gField[3];
  int frozenOn;           /* 1: material `frozen` is held still */
  float frozen;
} cco_sim_u;

/* A1 hashAt, src/glsl-sim.js:79-91 */
void cco_grid_make(int G, cco_grid* g);
void cco_hash_at(const cco_grid* g, int cx, int cy, int cz, int* ax, int* ay);
/* A2/A3 passSlots + seat codec, src/glsl-sim.js:103-156, src/nbody.js:379-399 */
void cco_slots(const float* pos, int N, float invCell, const cco_grid* g, uint32_t* seatA, uint32_t* seatB);
/* A8 boxOf, src/glsl-sim.js:25-37 */
int cco_box_of(const cco_box* b, const float p[3], float g[3]);
/* A4 DEP_VS/DEP_FS, src/glsl-sim.js:168-219 (accumulated in primitive order) */
void cco_deposit(const float* pos, const float* mat, int N, const cco_box* b, int box, const float mono[4],
                 float eps2, float* fine);
/* A5 COARSE_FS / MONO_FS, src/glsl-sim.js:222-251 */
void cco_coarse(const float* fine, float* coarse);
void cco_mono(const float* coarse, float* mono);
/* A6 FAR_FS, src/glsl-sim.js:263-295 */
void cco_far(const float* coarse, float* farA, float* farT0, float* farT1);
/* A7 CELL_FS, src/glsl-sim.js:303-352; eps2 = 0.5 FH^2, gm = G m */
void cco_cell(const float* fine, const float* coarse, const float* farA, const float* farT0, const float* farT1,
              float eps2, float gm, float* force);
/* A7 the same per-cell field in FP64 by brute force (the pinning check) */
void cco_cell_f64(const float* fine, const float* coarse, const float* farA, const float* farT0, const float* farT1,
                  double eps2, double gm, double* force);
/* A11 PPGATHER_FS / PPFORCE_FS, src/glsl-sim.js:676-754 */
void cco_pp_gather(const float* pos, const float* mat, const float* index, const cco_box* b, float* list, float* box);
void cco_pp_force(const float* list, const float* box, int count, float gm, float cut, float a2, const float* tab,
                  float* force);
/* A12 simFS + passSim, src/glsl-sim.js:362-533 */
void cco_sim(const cco_sim_u* u, const float* pos, const float* vel, const float* mat, const float* q,
             const uint32_t* seatA, const uint32_t* seatB, const float* force0, const float* mono0,
             const float* force1, const float* mono1, const float* ppSlot, const float* ppForce,
             float* oPos, float* oVel, float* oAux, float* oQ);
/* A14 RIGID_FS, src/glsl-sim.js:537-570 (rot: column-major mat3) */
void cco_rigid(int N, const float* pos, const float* vel, const float* aux, const float* q, const float bulk0[3],
               const float bulk1[3], const float cen0[3], const float cen1[3], const float spin0[3], const float spin1[3],
               const float rot0[9], const float rot1[9], float dt, float* oPos, float* oVel, float* oAux, float* oQ);
/* A15 PERM_FS, src/glsl-sim.js:578-602 */
void cco_perm(int N, const float* perm, const float* pos, const float* vel, const float* aux, const float* q,
              float* oPos, float* oVel, float* oAux, float* oQ);
/* A20 COUPLE_FS, chicxulub.html:255-315: args = centre[3], vimp, aimp, c1, invMu, pexp, stall, z, heat, tK,
 * frozen, ambient, melt (15 floats) */
void cco_couple(int N, const float* args, const float* pos, const float* vel, const float* aux, const float* mat,
                float* oPos, float* oVel, float* oAux, float* oQ);
/* A16 DIAG_FS, src/glsl-sim.js:608-667 */
void cco_diag(const cco_sim_u* u, const float* pos, const float* q, const uint32_t* seatA, const uint32_t* seatB,
              const float* force0, const float* mono0, const float* force1, const float* mono1, const float* ppSlot,
              const float* ppForce, float* out);
/* A15 Morton key + stable counting sort, src/nbody.js:871-887 (FP64 key arithmetic) */
void cco_morton(const float* pos, int N, const double core[3], int32_t* keys, float* perm, int32_t* into);
/* A17 books, src/nbody.js:831-850: out[0..13] = M, px, py, pz, Jx, Jy, Jz, KE, PE, Q, unseated, full, iron, drift */
void cco_books(int N, double mref, const float* pos, const float* vel, const float* mat, const float* diag, double* out);
/* FP64 direct sum of Newton gravity softened at a (a2 = a^2), x G m: out N*4 (acc, pot) */
void cco_direct_f64(int N, const float* pos, const float* mat, double gm, double a2, double* out);
/* src/math.js:8-15 Mulberry32 */
uint32_t cco_mulberry_next(uint32_t* state, double* out);
/* A10 meshPairTable, src/nbody.js:56-109 (bins x {force r^2, potential r}) */
void cco_mesh_pair_table(int bins, double cut, float* f, float* p);

/* friends-of-friends linking in the reference's order (src/nbody.js:647-684); returns the touching-pair count */
long cco_fof_link(int N, const float* P, double a, double touch, double spring, int32_t* root, int32_t* label,
                  double* e2out, long e2cap, double* elastic, int32_t* linkOff, int32_t* partner, long pcap);

#ifdef __cplusplus
}
#endif
#endif
