/* cc_hip.h -- C ABI of the MI355X (gfx950) simulation step of cosmic-collisions.
 *
 * The reference drives its step as WebGL2 passes from the engine object
 * CC.nbody (src/nbody.js:1095-1118).  This library is what that engine's
 * GL layer (src/gl.js + src/glsl-sim.js) becomes: every pass is an entry point
 * here, and a host engine (the JavaScript engine object in
 * cosmic-collisions_amd/js/, over the N-API addon) keeps the reference's
 * schedule, body setup and read-out around it.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative CC_E* code; the
 *    message is cc_last_error(ctx) (ctx may be NULL for create-time errors).
 *  - Host arrays are owned by the caller and copied; device memory by the ctx.
 *  - Particle arrays use the reference's texture order (index i = y*W + x):
 *    pos/vel/aux float[N*4], q/mat float[N].  Scalars that the reference
 *    computes in JavaScript numbers are passed as double and rounded to
 *    float exactly where WebGL's uniform upload rounds them.
 *  - Calls on one ctx are not thread-safe; all work is queued on the ctx's
 *    stream, and functions that return data wait for it.
 */
#ifndef CC_HIP_H
#define CC_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define CC_OK 0
#define CC_EINVAL -1     /* bad argument or size */
#define CC_EDEVICE -2    /* HIP runtime error */
#define CC_ESTATE -3     /* call out of order */
#define CC_EHOST -4      /* host resources (memory, threads) */

typedef struct cc_ctx cc_ctx;

/* The run's constants: sim fields of build()/soften() (src/nbody.js:114-250,
 * :545-550) and the static uniforms setStaticUniforms() derives from them
 * (src/nbody.js:255-336). */
typedef struct {
  int N, W, H;             /* particles; state texture width/height (N = W*H) */
  int G;                   /* contact grid: 64 or 128 cells a side (src/nbody.js:209) */
  double m;                /* mass of a mantle particle (uGm, 1/uInvM) */
  double a;                /* particle radius (uTouch = 2a, uA2 = a^2) */
  double cell;             /* contact cell edge, 2.6a */
  double dt, k, c, creep, damp, tK, kappa;
  double cpIron;           /* CP_IRON / CP_ROCK */
  double seatSlack;        /* SEAT_SLACK = 0.6 radii (src/nbody.js:41) */
  double spin0[3], spin1[3];
  int field;               /* 1: constant gravity gField (chicxulub.html:235) */
  double gField[3];
  int frozenOn;            /* 1: material `frozen` held still */
  double frozen;
  int meshEvery;           /* MESH_EVERY (src/nbody.js:538), 8 */
} cc_params;

/* The per-pass host values of passSim (src/nbody.js:458-484). */
typedef struct {
  int settle, kick;
  double gdt;
  double bulk0[3], bulk1[3], cen0[3], cen1[3];
} cc_sim_args;

/* src/nbody.js:1032-1092 init + build's allocations; replaces CC.gl's
 * simTarget/makeGrid/floatTex (src/gl.js:29-83) for the physics. */
int cc_create(const cc_params* p, cc_ctx** out);
int cc_destroy(cc_ctx* ctx);
const char* cc_last_error(cc_ctx* ctx);
/* setStaticUniforms after soften() (src/nbody.js:545-550, :255-336) */
int cc_set_params(cc_ctx* ctx, const cc_params* p);
/* texImage2D/texSubImage2D of the state textures (src/nbody.js:217-219) */
int cc_upload_state(cc_ctx* ctx, const float* pos, const float* vel, const float* mat, const float* q);
/* readPixels of the state (src/nbody.js:987-990); any pointer may be NULL */
int cc_download_state(cc_ctx* ctx, float* pos, float* vel, float* aux, float* q, float* mat);
/* boxes() (src/nbody.js:342-350): uLA, uLB as computed there, uBoxB */
int cc_set_boxes(cc_ctx* ctx, const double LA[3], const double LB[3], int boxB);
/* placeBoxes' priming of the second box's mono texel (src/nbody.js:365-367); also any mesh texture by name */
int cc_set_mono(cc_ctx* ctx, int box, const float mono[4]);
/* passSlots (src/nbody.js:379-399, SLOT_VS/FS src/glsl-sim.js:118-156) */
int cc_slots(cc_ctx* ctx);
/* the two seat atlases in the reference's RGBA32F layout, AW*AH*4 words each */
int cc_read_seats(cc_ctx* ctx, uint32_t* seatA, uint32_t* seatB);
/* passDeposit / passCoarse / passCell / passMesh (src/nbody.js:402-455) */
int cc_deposit(cc_ctx* ctx, int box);
int cc_coarse(cc_ctx* ctx, int box);
int cc_cell(cc_ctx* ctx, int box);
int cc_mesh(cc_ctx* ctx);
/* mesh textures: which = 0 fine (513x512x4), 1 coarse (64x64x4), 2 mono (4),
 * 3 farA, 4 farT0, 5 farT1 (64x64x4), 6 force (513x512x4) */
int cc_read_mesh(cc_ctx* ctx, int box, int which, float* out);
int cc_write_mesh(cc_ctx* ctx, int box, int which, const float* in);
/* PP_TAB (src/nbody.js:109, :1080-1082): 64 x {force r^2, potential r} */
int cc_set_pp_table(cc_ctx* ctx, const float* tab);
/* setPP (src/nbody.js:957-966): the loose list (particle indices), count <= 8192; runs passPP if count > 0 */
int cc_set_pp(cc_ctx* ctx, const int32_t* list, int count);
/* passPP (src/nbody.js:511-528) */
int cc_pp(cc_ctx* ctx);
/* pp.list (8192x4), pp.box (8192), pp.force (4 slices x 8192 x 4, slice-major); any may be NULL */
int cc_read_pp(cc_ctx* ctx, float* list, float* box, float* force);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): the pairwise pass of
 * cc_set_pp + cc_pp on the caller's DEVICE arrays, with no context and no
 * allocation.  pos float[N*4], mat float[N], index int32[count] (entries in
 * [0, N); others are empty seats), tab float[64*2] (the pair table), force
 * float[4*8192*4] out (slice-major, as cc_read_pp); LA/LB/boxB the mesh boxes
 * (as cc_set_boxes); m, a as in cc_params; stream a hipStream_t (NULL: the
 * null stream).  Queued on the stream, not waited for.  count <= 8192, the
 * size of its fixed workspace (a static device array, one per GPU, kept until
 * the process ends); larger counts fail with CC_EINVAL.  Calls that overlap on
 * one GPU share that workspace and race.  Errors: cc_last_error(NULL). */
int cc_pp_noalloc(const float* pos, const float* mat, int N, const double LA[3], const double LB[3], int boxB,
                  const int32_t* index, int count, double m, double a, const float* tab, float* force, void* stream);
/* passSim (src/nbody.js:458-484): one step pass, state src -> dst, swap */
int cc_sim(cc_ctx* ctx, const cc_sim_args* a);
/* step (src/nbody.js:551-558): seats every 2nd step, mesh + pairwise every
 * meshEvery-th with gravity as one meshEvery*dt impulse; with field gravity
 * (the crater, chicxulub.html:779-782) no mesh and gravity every step.  box_L is the
 * boxes() uniform for the mesh step (NULL: keep the last); stepNo advances. */
int cc_step(cc_ctx* ctx, int settle, int kick, const cc_sim_args* a);
/* n steps without settle or kick; boxL[n_mesh][7] = {LA, LB, boxB} for each mesh step in order (NULL: keep).
 * Whole cycles of meshEvery steps from a mesh step replay one captured HIP graph each (captured on first
 * use for the present params, loose-list count, box B and buffer parities); the rest run eagerly. */
int cc_step_n(cc_ctx* ctx, int n, const cc_sim_args* a, const double* boxL, int nBoxL);
/* step_n's cycle graphs on (default) or off (every step launched eagerly; same results) */
int cc_set_graphs(cc_ctx* ctx, int on);
/* One pass repeated back to back between HIP events on the ctx's stream, for
 * measurement: 0 the step pass (no settle, kick or gravity impulse: it
 * advances the state), 1 the seats, 2 the mesh (exact far field), 3 the
 * pairwise pass, 4 the mesh as on a mesh step between judges (the learned
 * far field if it is standing in), 5 the mesh as on a judge step, 6 a plain
 * copy of the positions (N x 16 B each way: the calibration of the HBM
 * traffic counters); *ms = device time per pass. */
int cc_time_pass(cc_ctx* ctx, int pass, int reps, double* ms);
int cc_get_step_no(cc_ctx* ctx, int* stepNo);
int cc_set_step_no(cc_ctx* ctx, int stepNo);
/* rigidStep's GPU half (RIGID_FS, src/glsl-sim.js:537-570): rot column-major */
int cc_rigid(cc_ctx* ctx, const double bulk0[3], const double bulk1[3], const double cen0[3], const double cen1[3],
             const double rot0[9], const double rot1[9], double dtF);
/* The crater's coupling at touchdown (COUPLE_FS + couple(), chicxulub.html:255-315, :820-842):
 * the excavation flow from the point source `centre`; MAT_MELT is `melt`. */
typedef struct {
  double centre[3];
  double vimp, aimp, c1, invMu, pexp, stall, z, heat, tK, frozen, ambient, melt;
} cc_couple_args;
int cc_couple(cc_ctx* ctx, const cc_couple_args* a);
/* DIAG_FS alone (src/glsl-sim.js:608-667) on the current seats/mesh/pp; out N*4 (NULL: keep on device) */
int cc_diag(cc_ctx* ctx, float* out);
/* Morton order (src/nbody.js:871-887) from the analysis' core centre, applied
 * to the state (PERM_FS + applyPerm, src/nbody.js:932-954) and the seats
 * rebuilt; perm[dst] = src and into[src] = dst returned (either may be NULL) */
int cc_morton(cc_ctx* ctx, const double core[3], int32_t* perm, int32_t* into);
/* the FP64 books over the state and the last DIAG (src/nbody.js:831-850):
 * out[13] = M, px, py, pz, Jx, Jy, Jz, KE, PE, Q, iron mass, unseated count, full count */
int cc_books(cc_ctx* ctx, double* out);

/* The analysis of the current state: fof() of src/nbody.js:646-870 without its Morton sort
 * (cc_morton does that from out->core).  Friends-of-friends groups at linking length 2.4 a
 * (2^20 hashed cells, union-find), the largest group's and the second's statistics, the
 * largest body's spin, the disk and the moon, the books and the loose list (at most cap
 * entries, farthest first past the cap).  Every field equals the reference's bit for bit:
 * the cells, the linked pairs and the components are found on the device; every FP64 sum
 * runs on the host in the reference's particle-index order. */
typedef struct {
  double m;           /* mref: mass of a unit of mat (sim.m) */
  double a;           /* particle radius (sim.a) */
  double spring;      /* contact stiffness for the elastic energy (sim.k) */
  double touch;       /* contact distance (2 a) */
  int32_t cap;        /* loose-list capacity (PP_CAP, at most 8192) */
  int32_t flags;      /* CC_ANALYSE_* */
  const float* diag;  /* N x 4 DIAG read-out to use, or NULL for the one cc_diag left on the device */
} cc_analyse_args;
/* replay the reference's union order on the host even when no tie among the largest
 * groups needs it (the replay is otherwise done only when it can change the result) */
#define CC_ANALYSE_REPLAY_UNIONS 1
typedef struct {      /* the second group, when it is a body apart (src/nbody.js:811-816) */
  double mass, dist;
  int32_t bound, pad_;
  double imp, iron, com[3], core[3], vel[3], R, Redge, Tsurf, hotCom[3], hotR;
} cc_second;
typedef struct cc_report {   /* fof()'s report (src/nbody.js:888-892), same names */
  double largest, largestImp, largestIron, day, tilt, axis[3], bodyVel[3], orbit, escape;
  double com[3], core[3], vel[3], R, Redge, Tsurf, hotCom[3], hotR, impMass, impCom[3], impRms;
  double M, iron, drift, L, KE, PE, EL, Q, unseated, full, jD, moon;
  int32_t hasSecond, pad_;
  cc_second second;
  int32_t ppCount;    /* entries written to list */
  int32_t ppLoose;    /* loose particles before the cap (pp.loose) */
  int32_t groups;     /* number of groups (no reference counterpart) */
  int32_t largestN, secondN;   /* member counts of the largest and second group */
  int32_t replayed;   /* 1 when the union order was replayed on the host */
  double ms[3];       /* device, transfer and host milliseconds of this call */
} cc_report;
int cc_analyse(cc_ctx* ctx, const cc_analyse_args* args, cc_report* out, int32_t* list);
/* analyze(false) and collect() (src/nbody.js:967-1025): the same analysis in
 * two halves.  cc_analyse_begin runs the device's part and reads the state back
 * (synchronously, as the reference's passDiag and read-back are issued), then
 * starts the report's FP64 sums on a host thread and returns; the caller steps
 * on meanwhile.  cc_analyse_end delivers that report: 0 with *out and list
 * filled, 1 (wait = 0) while the sums still run; wait != 0 waits for them.  One
 * analysis at a time: cc_analyse_begin while one is pending is CC_ESTATE, and a
 * synchronous cc_analyse drops a pending one (as the reference drops the
 * worker's result when it analyses synchronously). */
int cc_analyse_begin(cc_ctx* ctx, const cc_analyse_args* args);
int cc_analyse_end(cc_ctx* ctx, int wait, cc_report* out, int32_t* list);

/* wait for the ctx's stream */
int cc_sync(cc_ctx* ctx);
/* device time of the last cc_step_n call, ms (HIP events on the ctx's stream) */
int cc_last_step_ms(cc_ctx* ctx, double* ms);

/* The learned far field (no reference counterpart).  A 3D convolutional
 * operator on the 16^3 coarse mesh -- 4 levels, 32 channels, 3x3x3 kernels,
 * BF16 weights on the matrix cores with FP32 accumulation -- that stands in
 * for the block-to-block far field (FAR_FS, src/glsl-sim.js:263-295) on the
 * mesh steps between judges.  Every judgeEvery steps the exact far field runs,
 * the network's prediction is measured against it, and that exact pass becomes
 * the base; in between the network predicts each block's change of far-field
 * acceleration, potential and tidal tensor since the base from the change of
 * the coarse mesh (block masses and centre-of-mass offsets), added to the
 * base's far field and carried to each block's centre of mass by the tidal
 * tensor.  Trained on the GPU by the library's own forward and backward
 * kernels on pairs of passes recorded from the exact far field.
 * Weights: FP32, per layer [27 taps][cin][cout] then [cout] biases, layers in
 * order (cc_ff_weight_count floats). */
typedef struct cc_ff cc_ff;
int cc_ff_create(int capacity, cc_ff** out);     /* capacity: recorded passes, and training samples, on the device */
int cc_ff_destroy(cc_ff* f);
const char* cc_ff_last_error(cc_ff* f);
int cc_ff_weight_count(int64_t* n);
int cc_ff_get_weights(cc_ff* f, float* w, int64_t n);
int cc_ff_set_weights(cc_ff* f, const float* w, int64_t n);
/* the leaky ReLU's negative slope, 0.25 by default; 1 makes the network a linear operator of its input */
int cc_ff_set_slope(cc_ff* f, double slope);
/* record a pass: box `box` of ctx's last mesh pass, whose far field must be the exact one (cc_mesh, or a
 * cc_step_n without a network attached).  feat[4096][4] per block (voxel (z*16 + y)*16 + x): mass over the
 * box's total mass times 64, and that times the centre-of-mass offset from the block's centre in block
 * lengths; meta[4]: the box's corner offset L and total mass; record[4096][12]: the far field carried to
 * each block's centre (acceleration, potential, tidal xx yy zz xy xz yz, 2 spare) on occupied blocks and
 * their neighbours */
int cc_ff_record(cc_ff* f, cc_ctx* ctx, int box);
int cc_ff_passes(cc_ff* f, int* count);
int cc_ff_read_pass(cc_ff* f, int k, float* feat, float* meta, float* record);
int cc_ff_clear_passes(cc_ff* f);
/* a training sample from recorded passes: the prediction from `base` to `cur` as a mesh step forms it */
int cc_ff_pair(cc_ff* f, int base, int cur, int* sample);
/* samples in the network's own units: in[4096][16] (0-3 the present features, 4-7 their change since the
 * base over s -- the change's root mean square over blocks occupied in either pass --, 8-9 occupancy now
 * and at the base), target[4096][16] (the change of the record's first ten channels over s times the
 * total mass / 64, per block length squared, to the first and cubed), mask[4096] (1: occupied now) */
int cc_ff_push_sample(cc_ff* f, const float* in, const float* target, const float* mask);
int cc_ff_read_sample(cc_ff* f, int k, float* in, float* target, float* mask);
int cc_ff_samples(cc_ff* f, int* count);
int cc_ff_clear(cc_ff* f);
/* iters Adam steps (lr) on seeded batches (<= 64) of the record; losses[iters] */
int cc_ff_train(cc_ff* f, int iters, int batch, double lr, uint32_t seed, double* losses);
/* the network on samples [first, first+count): outputs (count*4096*16, or NULL) and the relative L2 error of the
 * accelerations and of all ten channels over occupied blocks */
int cc_ff_eval(cc_ff* f, int first, int count, float* out, double* relAcc, double* relAll);
/* the masked mean-squared loss and its gradient wrt every weight (no update), samples [first, first+count) */
int cc_ff_gradient(cc_ff* f, int first, int count, float* grad, double* loss);
/* attach (f) or detach (NULL) the network to a ctx's step: every judgeEvery steps (a multiple of meshEvery)
 * the exact far field makes the force mesh, the network's prediction is measured against it (the relative
 * error of the force mesh's far-field part over the whole force mesh), and that pass becomes the base; the
 * network stands in on the mesh steps in between while the last error was <= tol and measured from a base
 * that could serve (recorded since the state was last uploaded, the box moved by under a quarter block) */
int cc_set_far_field(cc_ctx* ctx, cc_ff* f, int judgeEvery, double tol);
/* per box (out[16]): last relative force error, judge passes, fallbacks, mesh steps run on the network, the
 * largest and the summed error and the count of judges measured from a base that could serve, spare */
int cc_ff_stats(cc_ctx* ctx, double* out);

/* Batched sweeps (no reference counterpart: the reference runs one instance
 * per page).  Independent instances, each its own ctx (any N = W*H up to
 * 262144), stepped together: a cycle of meshEvery steps of every instance is
 * one captured HIP graph, so the batch advances a cycle per launch.  Inside
 * it each step's step pass and seat pass are one kernel launch each for all
 * instances (a work-group map to per-instance jobs); the mesh step's own
 * passes run per instance, side by side.  Steps outside a whole cycle run
 * every instance on its own stream.  Results are those of each instance
 * stepped alone with cc_step_n.  The ctxs stay owned by the caller and must
 * outlive the batch; all must share meshEvery and stand at the same step of
 * the mesh cycle. */
typedef struct cc_batch cc_batch;
int cc_batch_create(cc_ctx* const* ctxs, int n, cc_batch** out);
int cc_batch_destroy(cc_batch* b);
const char* cc_batch_last_error(cc_batch* b);
/* n steps of every instance; args[n_inst]; boxL: the instances' mesh-step
 * boxes one after another ({LA, LB, boxB} x nBoxL[i] for instance i), or NULL */
int cc_batch_step_n(cc_batch* b, int n, const cc_sim_args* args, const double* boxL, const int* nBoxL);
/* device time of the last cc_batch_step_n, ms */
int cc_batch_last_ms(cc_batch* b, double* ms);

/* Launch shapes that cannot change a result (no reference counterpart): the
 * cell pass's work-groups (512, 1024, 2048 or 4096 striding over the occupied
 * texels; each texel's sum is its own), the step pass's shape (0: one
 * wavefront per 64 particles, consecutive work-groups spread over the XCDs;
 * 1: the same in launch order; 2: spread over the XCDs, three wavefronts per
 * 64 particles, one per z-plane of the neighbourhood, their candidate lists
 * joined in order -- for states too small to fill the GPU; 3: nine
 * wavefronts, one per row of three cells -- while the work-groups fit the
 * CUs at once; each particle's update is its own) and where a seat
 * rebuild's count is taken
 * (fuseCount 1: by the step pass before it, on the positions that pass
 * writes; 0: by a pass of its own at the rebuild; the same counts).  cc_autotune times every candidate on the
 * present state and keeps the fastest (a candidate must beat the current
 * choice by 2 %, the launch order by 5 %: it fetches about four times the
 * HBM bytes), the state as it was; it is called where the next step is a
 * mesh step, without constant gravity or a learned far field attached
 * (CC_ESTATE otherwise); it leaves fuseCount as it is.  cc_set_tuning sets all
 * three outright.  Defaults 2048, shape 3 up to 16,384 particles, 2 up to
 * 98,304 and 0 above, and 1.  Captured step cycles are
 * re-captured for a new choice. */
typedef struct {
  int32_t cellBlocks, simOrder;   /* the choice */
  double cellMs[4];               /* mesh pass per cell grid (512, 1024, 2048, 4096), ms */
  double simMs[4];                /* step pass per shape (0-3), ms */
} cc_tune_report;
int cc_autotune(cc_ctx* ctx, int reps, cc_tune_report* out);
int cc_set_tuning(cc_ctx* ctx, int cellBlocks, int simOrder, int fuseCount);

/* The order the step pass takes the particles in (no reference counterpart;
 * like the tuning, it cannot change a result: each particle's sums depend only
 * on the seats, which name particles by index).  The library sorts the
 * particles by the Morton code of their contact cell at the start of a mesh
 * cycle once `every` steps have passed since the last sort (and after an
 * upload of positions or a Morton permutation), so that a wavefront's lanes
 * share their neighbourhoods' cells and the caches hold them; every = 0 takes
 * the particles in index order.  Default 256. */
int cc_set_visit(cc_ctx* ctx, int every);

/* Body setup on the host (no device work): relax() of src/bodies.js:97-146,
 * the overlap relaxation of a freshly laid pile (Gauss-Seidel over a cell grid
 * rebuilt every fourth sweep, pairs closer than 0.96 x 2a pushed apart by half
 * their overlap), float64 in the reference's order on FP32 positions, so the
 * result is the reference's bit for bit.  The piles of a call are spread over
 * up to one thread per hardware thread (a pile is sequential in itself).  A
 * pile is held inside the spheroid of semi-axes A (x, z) and C (y) (the bodies,
 * src/bodies.js:74), or, CC_PILE_GROUND, inside radius A of the y axis and
 * between y = -C and y = 0 (the crater's ground patch and its confine
 * callback, chicxulub.html:543-548).  sweeps <= 0: the reference's 10. */
#define CC_PILE_SPHEROID 0
#define CC_PILE_GROUND 1
typedef struct {
  float* p;            /* n x 3 positions, relaxed in place */
  int32_t n, sweeps;
  double a, A, C;      /* particle radius, the confinement's two sizes */
  int32_t confine, pad_;
} cc_pile;
typedef struct {
  int32_t sweeps, pad_;
  double worst;        /* the last sweep's largest overlap over a */
} cc_pile_report;
int cc_relax_piles(const cc_pile* piles, int count, cc_pile_report* reports);
const char* cc_bodies_last_error(void);

/* The crater view's readout (craterStats, chicxulub.html:848-919), on the
 * host from a read-out state (pos and vel N x 4, R the masses in units of
 * mref): float64 in the reference's order, the same names.  Particles with
 * pos.w < 0.5 (the frozen shell) are left out; the settled surface is binned
 * in 96 rings out to Rbox for the rim and the crater's edge. */
typedef struct {
  double Rbox;             /* the ground patch's radius */
  double lSI, vSI, gSI;    /* metres, metres per second per length and speed unit; gravity m/s^2 */
  double rhoSI;            /* the target's density, kg/m^3 */
  double vOrb, vEsc;       /* orbital and escape speeds in speed units */
} cc_crater_args;
typedef struct {
  double craterD, rimR, rimH, depth;
  double excavated, thrown, platform, airborne, fastOrb, fastEsc, melt, vapour, Tmax;
  double mass, KE, PE, drift, Tsurf, R;
  double com[3];
} cc_crater_report;
int cc_crater_stats(int N, double mref, const float* pos, const float* vel, const float* R, const cc_crater_args* args,
                    cc_crater_report* out);

#ifdef __cplusplus
}
#endif
#endif
